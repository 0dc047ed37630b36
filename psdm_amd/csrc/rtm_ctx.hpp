// Context internals shared by the one-shot runtime (rtm_api.hip) and the batched
// engine (rtm_engine.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/psdm_rtm.h"

namespace psdm {
namespace host {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

struct SnapStore {
    float *d = nullptr;
    size_t bytes = 0;
    int ny = 0, nx = 0, nslots = 0;
};

}  // namespace host
}  // namespace psdm

struct psdm_rtm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, psdm::host::DevBuf> bufs;  // named workspace buffers, grown on demand
    std::map<long long, psdm::host::SnapStore> snaps;
    std::vector<std::pair<void *, size_t>> spare;  // released snapshot buffers
    int tile_x = 8;
    int stream_w = 2;  // interior columns per stream item of the one-shot solver
    int weights = 0;   // second-derivative weights of the device solver: 0 reference form, 1 zero-sum form
    int use_graph = 0, graph_steps = 64;
    int timing = 0;                                 // sample every `timing`-th step kernel (0 = off)
    std::vector<hipEvent_t> ev_free;                // recycled events
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
    long long n_timed = 0;
    double sum_ms = 0.0;
    // gather conditioning: filter taps of the last (nt_data, fe, typef, f_init, f_final);
    // with timing on, device time of the filter and resample kernels per call
    std::vector<unsigned> cond_key;
    hipEvent_t cond_ev[4] = {};
    double cond_ms[2] = {0.0, 0.0};
    long long cond_n[2] = {0, 0};
    // depth-to-time conversion: device time of its two kernels per call (timing on)
    hipEvent_t twt_ev[4] = {};
    double twt_ms[2] = {0.0, 0.0};
    long long twt_n = 0;
    // SU codec: device time of the decode [0] and encode [1] kernels per call (timing on)
    hipEvent_t su_ev[4] = {};
    double su_ms[2] = {0.0, 0.0};
    long long su_n[2] = {0, 0};
    void *su_pinned = nullptr;  // pinned staging of psdm_rtm_su_read_host
    size_t su_pinned_n = 0;
    // Kirchhoff: device time of the eikonal [0] and imaging [1] kernels per call (timing on)
    hipEvent_t kir_ev[4] = {};
    double kir_ms[2] = {0.0, 0.0};
    long long kir_n[2] = {0, 0};
    size_t eik_work_cap = (size_t)16 << 30;  // eikonal workspace limit (sources run in batches)
    std::map<unsigned, bool> div_ok;  // divisor bits -> its two-operation division matched the IEEE one
};

namespace psdm {
namespace host {

int fail(psdm_rtm_ctx *c, int code, const std::string &msg);
// Failed HIP call: records it and clears the thread's last-error state, so that a caller
// checking hipGetLastError after its own next launch does not see the library's error.
int hip_fail(psdm_rtm_ctx *c, hipError_t e, const char *expr);
// Grow-only named device buffer of the context.
int get_buf(psdm_rtm_ctx *c, const char *name, size_t bytes, void **out);
int set_device(psdm_rtm_ctx *c);
// 1D CPML damping profile (FP64 evaluation rounded to FP32), index = distance - 1.
void pml_profile(int dPML, double dt, double d, std::vector<float> &A, std::vector<float> &B);

}  // namespace host
}  // namespace psdm

#define HIP_TRY(ctx, expr)                                                                                   \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return psdm::host::hip_fail(ctx, e_, #expr);                                                     \
    } while (0)
