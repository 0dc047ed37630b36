"""MI355X-native RTM hot path of PSDM (2D acoustic FD + CPML, back-propagation,
zero-lag imaging) behind the reference's solver interface.

psdm_amd.solvers  -- the reference's solver entry points (solver_ac_*_ram/_disk),
                     dynamic_window, Ricker, and the device-resident form
psdm_amd._native  -- ctypes binding of libpsdm_rtm.so (include/psdm_rtm.h)
"""
