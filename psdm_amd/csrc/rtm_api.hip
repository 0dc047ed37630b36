// Host runtime of libpsdm_rtm.so: context, device buffers, the device-resident time
// loop and the C ABI declared in include/psdm_rtm.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "rtm_core.hpp"
#include "rtm_ctx.hpp"

using psdm::StepArgs;

namespace {

// Geometry of one call, 0-based padded-grid quantities.
struct Layout {
    int ny, nx, nt, dn, nyt, nxt, ld, lds, off, fs, ram, ss, nt_snap;
    float dx, dt;
    int tile_x, per_xcd;
    psdm::Tiling T;
};

}  // namespace

namespace psdm {
namespace host {

int fail(psdm_rtm_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(psdm_rtm_ctx *c, hipError_t e, const char *expr) {
    (void)hipGetLastError();
    return fail(c, PSDM_RTM_EDEVICE, std::string(expr) + ": " + hipGetErrorString(e));
}

int get_buf(psdm_rtm_ctx *c, const char *name, size_t bytes, void **out) {
    DevBuf &b = c->bufs[name];
    if (b.n < bytes) {
        // queued kernels may still read the old buffer
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (b.p) HIP_TRY(c, hipFree(b.p));
        b.p = nullptr;
        b.n = 0;
        size_t want = std::max<size_t>(bytes, 256);
        HIP_TRY(c, hipMalloc(&b.p, want));
        b.n = want;
    }
    *out = b.p;
    return 0;
}

int set_device(psdm_rtm_ctx *c) {
    HIP_TRY(c, hipSetDevice(c->device));
    return 0;
}

// 1D CPML damping profile, PML_coeffs (solver_PML_coeffs.f90:16-34), evaluated in FP64
// and rounded to FP32.  Index = distance into the CPML - 1 (interior side first);
// the first acc=3 cells keep A = B = 0 (:272 starts at i = 1+acc).
void pml_profile(int dPML, double dt, double d, std::vector<float> &A, std::vector<float> &B) {
    const int acc = 3, dn = dPML + 2 * acc;
    std::vector<double> a(dn, 0.0), b(dn, 0.0);
    const double pi = 3.14159265, Refl = 0.00001, w = (double)(dPML - 1) * d;
    for (int i = acc; i < dPML + acc; i++) {
        const double y = (double)(i - acc) * d;
        const double alpha = 30. * pi * (1. - (y / w));
        const double sigma = (-(6000. * std::log(Refl)) / (2. * w)) * ((y / w) * (y / w));
        a[i] = std::exp(-(sigma + alpha) * dt);
        b[i] = (a[i] - 1.) * sigma / (sigma + alpha);
    }
    for (int i = dPML + acc; i < dn; i++) {
        a[i] = a[dPML + acc - 1];
        b[i] = b[dPML + acc - 1];
    }
    A.resize(dn);
    B.resize(dn);
    for (int i = 0; i < dn; i++) {
        A[i] = (float)a[i];
        B[i] = (float)b[i];
    }
}

}  // namespace host
}  // namespace psdm

using namespace psdm::host;

namespace {

// Snapshot buffers are recycled: a released store goes to the spare list (kept in
// stream order, so a later forward may reuse it without waiting) instead of hipFree,
// which would stall the stream between shots.
int free_store(psdm_rtm_ctx *c, long long key) {
    auto it = c->snaps.find(key);
    if (it == c->snaps.end()) return 0;
    if (it->second.d) c->spare.emplace_back(it->second.d, it->second.bytes);
    c->snaps.erase(it);
    return 0;
}

int alloc_store(psdm_rtm_ctx *c, size_t bytes, float **out) {
    *out = nullptr;
    if (bytes == 0) return 0;
    // best fit among spares
    int best = -1;
    for (size_t q = 0; q < c->spare.size(); q++)
        if (c->spare[q].second >= bytes && (best < 0 || c->spare[q].second < c->spare[best].second)) best = (int)q;
    if (best >= 0) {
        *out = (float *)c->spare[best].first;
        c->spare.erase(c->spare.begin() + best);
        return 0;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto &sp : c->spare) HIP_TRY(c, hipFree(sp.first));  // too small: release them
    c->spare.clear();
    void *p = nullptr;
    HIP_TRY(c, hipMalloc(&p, bytes));
    *out = (float *)p;
    return 0;
}

int check_geom(psdm_rtm_ctx *c, const psdm_rtm_geom *g) {
    if (!g) return fail(c, PSDM_RTM_EINVAL, "null geometry");
    if (g->ny < 1 || g->nx < 1) return fail(c, PSDM_RTM_EINVAL, "ny and nx must be >= 1");
    if (g->nt < 2) return fail(c, PSDM_RTM_EINVAL, "nt must be >= 2");
    if (g->dPML < 2 || g->dPML + 6 > psdm::kMaxTables) return fail(c, PSDM_RTM_EINVAL, "dPML out of range");
    if (g->store_snap < 1) return fail(c, PSDM_RTM_EINVAL, "store_snap must be >= 1");
    if (g->nt_snap < 0) return fail(c, PSDM_RTM_EINVAL, "nt_snap must be >= 0");
    if (!(g->dt > 0.f) || !(g->dmodel > 0.f)) return fail(c, PSDM_RTM_EINVAL, "dt and dmodel must be > 0");
    if (g->variant == 0 && g->nx < g->dPML + 6)
        return fail(c, PSDM_RTM_EINVAL,
                    "the _ram padding loops (solver_ac_forward_ram.f90:288-302) read ctmp(1,j) for j up to "
                    "dPML+6: nx must be >= dPML+6");
    if (g->NumRec < 0 || g->NumSou < 0) return fail(c, PSDM_RTM_EINVAL, "negative point count");
    if (g->ny >= 65536 || g->nx >= 32768) return fail(c, PSDM_RTM_EINVAL, "window exceeds 65535 x 32767 cells");
    const long long nxt = g->nx + 2LL * (g->dPML + 6);
    if ((long long)psdm::state_ld(g->ny, g->dPML + 6) * nxt >= (1LL << 31))
        return fail(c, PSDM_RTM_EINVAL, "padded grid exceeds 2^31 cells");
    return 0;
}

Layout make_layout(const psdm_rtm_geom *g, int tile_x, int stream_w) {
    Layout L;
    L.ny = g->ny;
    L.nx = g->nx;
    L.nt = g->nt;
    L.dn = g->dPML + 6;
    L.nyt = g->ny + 2 * L.dn;
    L.nxt = g->nx + 2 * L.dn;
    L.ld = psdm::state_ld(L.ny, L.dn);
    L.lds = psdm::interior_ld(L.ny);
    L.off = psdm::state_off(L.dn);
    L.fs = g->fs != 0;
    L.ram = g->variant == 0;
    L.ss = g->store_snap;
    L.nt_snap = g->nt_snap;
    L.dx = g->dmodel;
    L.dt = g->dt;
    L.tile_x = tile_x;
    L.T = psdm::make_tiling(L.ny, L.nx, L.dn, tile_x, stream_w, L.ram != 0, L.fs != 0);  // state zeroed per call
    L.per_xcd = (L.T.nitems + 7) / 8;
    return L;
}

// State shared by the forward and backward loops.
struct Prepared {
    StepArgs a;
    float *p0, *p1, *phx0, *phx1, *phy0, *phy1, *hs0, *hs1;
};

// Allocates/zeroes the wavefield state, computes c2, CPML tables, injection events
// (last writer wins per cell, interior cells only: :348-358, :443-444) and receivers.
int prepare(psdm_rtm_ctx *c, const Layout &L, const float *d_ctmp, const float *d_traces, int ntr, const int *px,
            const int *py, int npts, const int *rx, const int *ry, int nrec, Prepared &P) {
    StepArgs &a = P.a;
    std::memset(&a, 0, sizeof(a));
    const size_t ld = L.ld, nxt = L.nxt, dn2 = 2 * L.dn;
    const size_t n_p = (ld * nxt + 32 + 63) / 64 * 64, n_x = ld * dn2, n_y = dn2 * nxt;  // n_p: + row-offset room
    void *v;
    // one allocation for the state that must start at zero
    const size_t n_state = 2 * n_p + 3 * n_x + 3 * n_y + 2 * dn2;
    if (int rc = get_buf(c, "state", n_state * sizeof(float), &v)) return rc;
    float *s = (float *)v;
    HIP_TRY(c, hipMemsetAsync(s, 0, n_state * sizeof(float), c->stream));
    P.p0 = s + L.off;
    P.p1 = s + n_p + L.off;
    P.phx0 = s + 2 * n_p;
    P.phx1 = P.phx0 + n_x;
    a.chx = P.phx1 + n_x;
    P.phy0 = a.chx + n_x;
    P.phy1 = P.phy0 + n_y;
    a.chy = P.phy1 + n_y;
    P.hs0 = a.chy + n_y;
    P.hs1 = P.hs0 + dn2;

    if (int rc = get_buf(c, "c2", ((size_t)L.lds * L.nx + 64) * sizeof(float), &v)) return rc;
    a.c2 = (float *)v + psdm::kIntPad;
    a.ldc = L.lds;
    psdm::launch_c2(d_ctmp, (float *)a.c2, L.ny, L.nx, L.lds, L.dt, L.dx, c->stream);

    std::vector<float> fA, fB;
    pml_profile(L.dn - 6, (double)L.dt, (double)L.dx, fA, fB);
    std::vector<float> tabs(fA);
    tabs.insert(tabs.end(), fB.begin(), fB.end());
    if (int rc = get_buf(c, "tabs", tabs.size() * sizeof(float), &v)) return rc;
    HIP_TRY(c, hipMemcpyAsync(v, tabs.data(), tabs.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    a.fA = (float *)v;
    a.fB = (float *)v + L.dn;

    // injection events, per stream item, ordered by column then row
    std::map<int, int> last;  // (iw << 16 | jw) -> trace column, last writer wins
    for (int s2 = 0; s2 < npts; s2++) {
        const int jw = py[s2] - 1, iw = px[s2] - 1;
        if (jw < 0 || jw >= L.ny || iw < 0 || iw >= L.nx) continue;  // only interior cells get source
        last[(iw << 16) | jw] = s2;
    }
    std::vector<std::vector<std::pair<int, int>>> per_item(L.T.nstream);
    for (auto &kv : last) per_item[psdm::stream_item_of_cell(L.T, kv.first & 0xffff, kv.first >> 16)].push_back(kv);
    std::vector<int> ev_ptr(L.T.nstream + 1, 0), ev_cell, ev_inj, cols;
    for (int u = 0; u < L.T.nstream; u++) {
        for (auto &kv : per_item[u]) {
            ev_cell.push_back(kv.first);
            ev_inj.push_back((int)cols.size());
            cols.push_back(kv.second);
        }
        ev_ptr[u + 1] = (int)ev_cell.size();
    }
    const int ninj = (int)cols.size();
    if (ninj > 0 && ntr <= 0) return fail(c, PSDM_RTM_EINVAL, "injection points without traces");
    std::vector<int> ints(ev_ptr);
    const size_t o_cell = ints.size();
    ints.insert(ints.end(), ev_cell.begin(), ev_cell.end());
    const size_t o_inj = ints.size();
    ints.insert(ints.end(), ev_inj.begin(), ev_inj.end());
    const size_t o_cols = ints.size();
    ints.insert(ints.end(), cols.begin(), cols.end());
    // receivers: cell j + i*ld of (dn+nyRec, dn+nxRec), or -1 outside the padded grid
    const size_t o_rec = ints.size();
    for (int r = 0; r < nrec; r++) {
        const int j = L.dn + ry[r] - 1, i = L.dn + rx[r] - 1;
        ints.push_back((j >= 0 && j < L.nyt && i >= 0 && i < L.nxt) ? j + i * L.ld : -1);
    }
    if (int rc = get_buf(c, "ints", ints.size() * sizeof(int), &v)) return rc;
    int *di = (int *)v;
    HIP_TRY(c, hipMemcpyAsync(di, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    a.ev_ptr = di;
    a.ev_cell = di + o_cell;
    a.ev_inj = di + o_inj;
    a.rec_cell = di + o_rec;
    a.nrec = nrec;
    a.ninj = ninj;
    if (ninj > 0) {
        if (int rc = get_buf(c, "inj", (size_t)(L.nt - 1) * ninj * sizeof(float), &v)) return rc;
        a.inj_val = (float *)v;
        psdm::launch_gather(d_traces, L.nt, di + o_cols, ninj, (float *)v, 0, c->stream);
    }
    // host vectors die here: wait for the copies before returning
    HIP_TRY(c, hipStreamSynchronize(c->stream));

    a.ny = L.ny;
    a.nx = L.nx;
    a.dn = L.dn;
    a.nyt = L.nyt;
    a.nxt = L.nxt;
    a.ld = L.ld;
    psdm::set_tiling(a, L.T);
    a.per_xcd = L.per_xcd;
    a.fs = L.fs;
    a.ram = L.ram;
    a.dt = L.dt;
    a.ss_f = (float)L.ss;
    a.zsum = c->weights == 1;
    a.lds = L.lds;
    a.snap_stride = (size_t)L.lds * L.nx;
    a.scale_stride = L.T.ri * L.nx;
    a.rec_row = -1;
    a.snap_slot = -1;
    return 0;
}

// Folds completed event pairs into the step-kernel statistics and recycles them.
int harvest_events(psdm_rtm_ctx *c) {
    for (auto &pr : c->ev_pending) {
        HIP_TRY(c, hipEventSynchronize(pr.second));
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, pr.first, pr.second));
        c->n_timed++;
        c->sum_ms += ms;
        c->ev_free.push_back(pr.first);
        c->ev_free.push_back(pr.second);
    }
    c->ev_pending.clear();
    return 0;
}

int take_event(psdm_rtm_ctx *c, hipEvent_t *e) {
    if (c->ev_free.empty() && c->ev_pending.size() >= 512)
        if (int rc = harvest_events(c)) return rc;
    if (c->ev_free.empty()) {
        HIP_TRY(c, hipEventCreate(e));
        return 0;
    }
    *e = c->ev_free.back();
    c->ev_free.pop_back();
    return 0;
}

int run_loop(psdm_rtm_ctx *c, const Layout &L, Prepared &P, int mode, int nslots_written, float *d_rec) {
    StepArgs a = P.a;
    a.mode = mode;
    float *pc = P.p0, *pn = P.p1, *phxo = P.phx0, *phxn = P.phx1, *phyo = P.phy0, *phyn = P.phy1;
    float *hsc = P.hs0, *hsn = P.hs1;
    const bool timing = c->timing > 0;
    a.rec_ld = L.nt;
    for (int k = 2; k <= L.nt; k++) {
        a.k = k;
        a.pc = pc;
        a.pn = pn;
        a.phx_o = phxo;
        a.phx_n = phxn;
        a.phy_o = phyo;
        a.phy_n = phyn;
        a.hs_c = hsc;
        a.hs_n = hsn;
        a.snap_slot = -1;
        if ((k - 1) % L.ss == 0) {
            const int s = (k - 1) / L.ss;  // forward: snap = 1, 2, ... (:478-480)
            if (mode == 0) {
                if (s <= nslots_written) a.snap_slot = s - 1;
            } else {
                const int sb = L.nt_snap - s;  // backward: snap counts down from nt_snap (:459-461)
                if (sb >= 1 && sb <= nslots_written) a.snap_slot = sb - 1;
            }
        }
        a.rec_row = (mode == 0 && d_rec && k >= 3) ? k - 3 : -1;
        a.rec = d_rec;
        const bool t_this = timing && (k % c->timing == 0);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (t_this) {
            if (int rc = take_event(c, &e0)) return rc;
            if (int rc = take_event(c, &e1)) return rc;
            c->ev_pending.emplace_back(e0, e1);
        }
        psdm::launch_step(a, L.tile_x, c->stream, e0, e1);
        std::swap(pc, pn);
        std::swap(phxo, phxn);
        std::swap(phyo, phyn);
        std::swap(hsc, hsn);
    }
    HIP_TRY(c, hipGetLastError());
    if (mode == 0 && d_rec && a.nrec > 0) {
        // traces of the last step, then Rec(nt) = Rec(nt-1) (:575-577)
        psdm::launch_record(pc, a.rec_cell, d_rec, L.nt, L.nt - 2, a.nrec, c->stream);
        HIP_TRY(c, hipMemcpy2DAsync(d_rec + (L.nt - 1), (size_t)L.nt * sizeof(float), d_rec + (L.nt - 2),
                                    (size_t)L.nt * sizeof(float), sizeof(float), a.nrec, hipMemcpyDeviceToDevice,
                                    c->stream));
    }
    return 0;
}

// Forward/backward on device.  Traces go to d_rec, Rec_point(nt, NumRec) column-major.
int fwd_device(psdm_rtm_ctx *c, const psdm_rtm_geom *g, const float *d_ctmp, const float *d_S, float *d_rec,
               long long key) {
    if (int rc = check_geom(c, g)) return rc;
    const Layout L = make_layout(g, c->tile_x, c->stream_w);
    Prepared P;
    if (int rc = prepare(c, L, d_ctmp, d_S, g->NumSou, g->nxSou, g->nySou, g->NumSou, g->nxRec, g->nyRec, g->NumRec, P))
        return rc;
    // snapshots written by this forward: s = 1 .. min(nt_snap, (nt-1)/store_snap)
    const int nslots = std::min(L.nt_snap, (L.nt - 1) / L.ss);
    if (int rc = free_store(c, key)) return rc;
    SnapStore st;
    st.ny = L.ny;
    st.nx = L.nx;
    st.nslots = nslots;
    st.bytes = nslots > 0 ? ((size_t)nslots * L.lds * L.nx + 64) * sizeof(float) : 0;
    if (int rc = alloc_store(c, st.bytes, &st.d)) return rc;
    c->snaps[key] = st;
    P.a.snaps = st.d ? st.d + psdm::kIntPad : nullptr;
    return run_loop(c, L, P, 0, nslots, d_rec);
}

int bwd_device(psdm_rtm_ctx *c, const psdm_rtm_geom *g, const float *d_ctmp, const float *d_Sadj, float *d_kernel,
               long long key) {
    if (int rc = check_geom(c, g)) return rc;
    auto it = c->snaps.find(key);
    if (it == c->snaps.end()) return fail(c, PSDM_RTM_ENOENT, "no device-resident snapshots under this key");
    if (it->second.ny != g->ny || it->second.nx != g->nx)
        return fail(c, PSDM_RTM_EINVAL, "snapshot window differs from the backward window");
    const Layout L = make_layout(g, c->tile_x, c->stream_w);
    Prepared P;
    if (int rc = prepare(c, L, d_ctmp, d_Sadj, g->NumRec, g->nxRec, g->nyRec, g->NumRec, nullptr, nullptr, 0, P))
        return rc;
    P.a.snaps = it->second.d ? it->second.d + psdm::kIntPad : nullptr;
    // the kernel accumulates into an image with leading dimension lds (kIntPad front
    // rows), seeded from the caller's kernel(ny,nx) and copied back afterwards
    void *v;
    if (int rc = get_buf(c, "img", ((size_t)L.lds * L.nx + 64) * sizeof(float), &v)) return rc;
    float *img = (float *)v + psdm::kIntPad;
    HIP_TRY(c, hipMemcpy2DAsync(img, (size_t)L.lds * sizeof(float), d_kernel, (size_t)L.ny * sizeof(float),
                                (size_t)L.ny * sizeof(float), L.nx, hipMemcpyDeviceToDevice, c->stream));
    P.a.image = img;
    if (int rc = run_loop(c, L, P, 1, it->second.nslots, nullptr)) return rc;
    HIP_TRY(c, hipMemcpy2DAsync(d_kernel, (size_t)L.ny * sizeof(float), img, (size_t)L.lds * sizeof(float),
                                (size_t)L.ny * sizeof(float), L.nx, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return free_store(c, key);
}

psdm_rtm_geom host_geom(int dPML, int fs, int variant, int NumSou, int NumRec, const int *nxSou, const int *nySou,
                        const int *nxRec, const int *nyRec, float dmodel, float dt, int ny, int nx, int store_snap,
                        int nt_snap, int nt) {
    psdm_rtm_geom g;
    g.ny = ny;
    g.nx = nx;
    g.nt = nt;
    g.dPML = dPML;
    g.fs = fs;
    g.variant = variant;
    g.store_snap = store_snap;
    g.nt_snap = nt_snap;
    g.dmodel = dmodel;
    g.dt = dt;
    g.NumSou = NumSou;
    g.NumRec = NumRec;
    g.nxSou = nxSou;
    g.nySou = nySou;
    g.nxRec = nxRec;
    g.nyRec = nyRec;
    return g;
}

// Host-array forward: upload, run, download traces (and snapshots when asked).
int host_forward(psdm_rtm_ctx *c, const psdm_rtm_geom &g, const float *S, const float *ctmp, float *Fields,
                 int host_fields, float *Rec_point, long long key) {
    if (int rc = set_device(c)) return rc;
    if (int rc = check_geom(c, &g)) return rc;
    if (g.NumSou > 0 && !S) return fail(c, PSDM_RTM_EINVAL, "null source array");
    if (g.NumRec > 0 && !Rec_point) return fail(c, PSDM_RTM_EINVAL, "null Rec_point");
    void *v;
    const size_t nv = (size_t)g.ny * g.nx, ns = (size_t)g.nt * std::max(g.NumSou, 1),
                 nr = (size_t)g.nt * std::max(g.NumRec, 1);
    if (int rc = get_buf(c, "h_io", (nv + ns + nr) * sizeof(float), &v)) return rc;
    float *d_vel = (float *)v, *d_S = d_vel + nv, *d_rec = d_S + ns;
    HIP_TRY(c, hipMemcpyAsync(d_vel, ctmp, nv * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (g.NumSou > 0)
        HIP_TRY(c, hipMemcpyAsync(d_S, S, (size_t)g.nt * g.NumSou * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (int rc = fwd_device(c, &g, d_vel, d_S, d_rec, key)) return rc;
    if (g.NumRec > 0)
        HIP_TRY(c, hipMemcpyAsync(Rec_point, d_rec, (size_t)g.nt * g.NumRec * sizeof(float), hipMemcpyDeviceToHost,
                                  c->stream));
    if (host_fields) {
        SnapStore &st = c->snaps[key];
        if (st.nslots > 0) {
            if (!Fields) return fail(c, PSDM_RTM_EINVAL, "null Fields with host_fields");
            // Fields(ny, nx, nt_snap) <- snapshots with leading dimension lds
            HIP_TRY(c, hipMemcpy2DAsync(Fields, (size_t)g.ny * sizeof(float), st.d + psdm::kIntPad,
                                        (size_t)psdm::interior_ld(g.ny) * sizeof(float), (size_t)g.ny * sizeof(float),
                                        (size_t)st.nslots * g.nx, hipMemcpyDeviceToHost, c->stream));
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return free_store(c, key);
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

int host_backward(psdm_rtm_ctx *c, const psdm_rtm_geom &g, const float *Sadj, const float *ctmp,
                  const float *Fields, int host_fields, float *kernel, long long key) {
    if (int rc = set_device(c)) return rc;
    if (int rc = check_geom(c, &g)) return rc;
    if (!kernel) return fail(c, PSDM_RTM_EINVAL, "null kernel");
    if (g.NumRec > 0 && !Sadj) return fail(c, PSDM_RTM_EINVAL, "null Sadj");
    void *v;
    const size_t nv = (size_t)g.ny * g.nx, ns = (size_t)g.nt * std::max(g.NumRec, 1);
    if (int rc = get_buf(c, "h_io", (nv + ns + nv) * sizeof(float), &v)) return rc;
    float *d_vel = (float *)v, *d_S = d_vel + nv, *d_img = d_S + ns;
    HIP_TRY(c, hipMemcpyAsync(d_vel, ctmp, nv * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (g.NumRec > 0)
        HIP_TRY(c, hipMemcpyAsync(d_S, Sadj, (size_t)g.nt * g.NumRec * sizeof(float), hipMemcpyHostToDevice,
                                  c->stream));
    HIP_TRY(c, hipMemcpyAsync(d_img, kernel, nv * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (host_fields) {
        // snapshots come from the caller's Fields(ny,nx,nt_snap)
        if (int rc = free_store(c, key)) return rc;
        SnapStore st;
        st.ny = g.ny;
        st.nx = g.nx;
        st.nslots = g.nt_snap;
        if (st.nslots > 0) {
            if (!Fields) return fail(c, PSDM_RTM_EINVAL, "null Fields with host_fields");
            const size_t lds = psdm::interior_ld(g.ny);
            st.bytes = ((size_t)st.nslots * lds * g.nx + 64) * sizeof(float);
            if (int rc = alloc_store(c, st.bytes, &st.d)) return rc;
            c->snaps[key] = st;
            HIP_TRY(c, hipMemcpy2DAsync(st.d + psdm::kIntPad, lds * sizeof(float), Fields, (size_t)g.ny * sizeof(float),
                                        (size_t)g.ny * sizeof(float), (size_t)st.nslots * g.nx, hipMemcpyHostToDevice,
                                        c->stream));
        } else {
            c->snaps[key] = st;
        }
    }
    if (int rc = bwd_device(c, &g, d_vel, d_S, d_img, key)) return rc;
    HIP_TRY(c, hipMemcpyAsync(kernel, d_img, nv * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Keys: _ram snapshots are keyed by the caller's Fields address, _disk ones by shot
// id; the two ranges cannot collide (ids are tagged in the top bit).
long long disk_key(int shotid) { return (long long)(0x4000000000000000LL | (unsigned)shotid); }

}  // namespace

extern "C" {

int psdm_rtm_create(psdm_rtm_ctx **out, int device) {
    if (!out) return PSDM_RTM_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PSDM_RTM_EDEVICE;
    psdm_rtm_ctx *c = new psdm_rtm_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PSDM_RTM_EDEVICE;
    }
    *out = c;
    return 0;
}

int psdm_rtm_destroy(psdm_rtm_ctx *c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->snaps)
        if (kv.second.d) (void)hipFree(kv.second.d);
    for (auto &sp : c->spare) (void)hipFree(sp.first);
    for (auto &kv : c->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    for (auto &pr : c->ev_pending) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (auto e : c->ev_free) (void)hipEventDestroy(e);
    for (auto e : c->cond_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : c->twt_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : c->su_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->su_pinned) (void)hipHostFree(c->su_pinned);
    for (auto e : c->kir_ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

const char *psdm_rtm_last_error(const psdm_rtm_ctx *c) { return c ? c->err.c_str() : "null context"; }

int psdm_rtm_forward_ram(psdm_rtm_ctx *c, int dPML, int fs, int NumSou, int NumRec, const int *nxSou,
                         const int *nySou, const int *nxRec, const int *nyRec, const float *S, float dmodel,
                         float dt, int ny, int nx, const float *ctmp, int store_snap, int nt_snap, int nt,
                         float *Fields, int host_fields, float *Rec_point) {
    if (!c) return PSDM_RTM_EINVAL;
    const psdm_rtm_geom g = host_geom(dPML, fs, 0, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, dmodel, dt, ny, nx,
                                      store_snap, nt_snap, nt);
    return host_forward(c, g, S, ctmp, Fields, host_fields, Rec_point, (long long)(uintptr_t)Fields);
}

int psdm_rtm_backward_ram(psdm_rtm_ctx *c, int dPML, int fs, int NumSou, int NumRec, const int *nxSou,
                          const int *nySou, const int *nxRec, const int *nyRec, const float *Sadj, float dmodel,
                          float dt, int ny, int nx, const float *ctmp, int store_snap, int nt_snap, int nt,
                          const float *Fields, int host_fields, float *kernel) {
    if (!c) return PSDM_RTM_EINVAL;
    const psdm_rtm_geom g = host_geom(dPML, fs, 0, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, dmodel, dt, ny, nx,
                                      store_snap, nt_snap, nt);
    return host_backward(c, g, Sadj, ctmp, Fields, host_fields, kernel, (long long)(uintptr_t)Fields);
}

int psdm_rtm_forward_disk(psdm_rtm_ctx *c, int dPML, int fs, int NumSou, int NumRec, const int *nxSou,
                          const int *nySou, const int *nxRec, const int *nyRec, const float *S, float dmodel,
                          float dt, int ny, int nx, const float *ctmp, int shotid, int store_snap, int nt_snap,
                          int nt, float *Rec_point) {
    if (!c) return PSDM_RTM_EINVAL;
    const psdm_rtm_geom g = host_geom(dPML, fs, 1, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, dmodel, dt, ny, nx,
                                      store_snap, nt_snap, nt);
    return host_forward(c, g, S, ctmp, nullptr, 0, Rec_point, disk_key(shotid));
}

int psdm_rtm_backward_disk(psdm_rtm_ctx *c, int dPML, int fs, int NumSou, int NumRec, const int *nxSou,
                           const int *nySou, const int *nxRec, const int *nyRec, const float *Sadj, float dmodel,
                           float dt, int ny, int nx, const float *ctmp, int shotid, int store_snap, int nt_snap,
                           int nt, float *kernel) {
    if (!c) return PSDM_RTM_EINVAL;
    const psdm_rtm_geom g = host_geom(dPML, fs, 1, NumSou, NumRec, nxSou, nySou, nxRec, nyRec, dmodel, dt, ny, nx,
                                      store_snap, nt_snap, nt);
    return host_backward(c, g, Sadj, ctmp, nullptr, 0, kernel, disk_key(shotid));
}

int psdm_rtm_forward_device(psdm_rtm_ctx *c, const psdm_rtm_geom *g, const float *d_ctmp, const float *d_S,
                            float *d_rec, long long key) {
    if (!c) return PSDM_RTM_EINVAL;
    if (int rc = set_device(c)) return rc;
    if (int rc = check_geom(c, g)) return rc;
    if (g->NumRec > 0 && !d_rec) return fail(c, PSDM_RTM_EINVAL, "null d_rec");
    return fwd_device(c, g, d_ctmp, d_S, d_rec, key);
}

int psdm_rtm_backward_device(psdm_rtm_ctx *c, const psdm_rtm_geom *g, const float *d_ctmp, const float *d_Sadj,
                             float *d_kernel, long long key) {
    if (!c) return PSDM_RTM_EINVAL;
    if (int rc = set_device(c)) return rc;
    if (!d_kernel) return fail(c, PSDM_RTM_EINVAL, "null d_kernel");
    return bwd_device(c, g, d_ctmp, d_Sadj, d_kernel, key);
}

int psdm_rtm_stack_image(psdm_rtm_ctx *c, float *d_kernel, int ny, int nx_w, int x_i, double *d_psdm_sum,
                         int *d_cov_sum, int nx_prop) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!d_kernel || !d_psdm_sum || !d_cov_sum) return fail(c, PSDM_RTM_EINVAL, "null pointer");
    if (ny < 1 || nx_w < 1 || x_i < 1 || x_i + nx_w - 1 > nx_prop)
        return fail(c, PSDM_RTM_EINVAL, "window [x_i, x_i+nx_w-1] outside 1..nx_prop");
    if (int rc = set_device(c)) return rc;
    psdm::launch_stack(d_kernel, nullptr, ny, ny, nx_w, x_i - 1, d_psdm_sum, ny, d_cov_sum, c->stream);
    HIP_TRY(c, hipGetLastError());
    return 0;
}

int psdm_rtm_taper_snapshots(psdm_rtm_ctx *c, const float *Fields, int nt_prop2, int N_taper) {
    if (!c) return PSDM_RTM_EINVAL;
    if (nt_prop2 < 1 || N_taper < 1) return fail(c, PSDM_RTM_EINVAL, "taper: need nt_prop2 >= 1 and N_taper >= 1");
    auto it = c->snaps.find((long long)(uintptr_t)Fields);
    if (it == c->snaps.end()) return fail(c, PSDM_RTM_ENOENT, "no device-resident snapshots under this key");
    if (int rc = set_device(c)) return rc;
    const SnapStore &st = it->second;
    // taper_pad_trace: i0 = max(1, nt_prop2 - N_taper + 1); trace(i0+n) *= w(n) for
    // i0+n <= nt_prop2, w(n) = 0.5*(1 + cos(pi*n/N_taper)), pi = 4 atan(1), FP32; snapshot
    // s (1-based) is slot s-1 of the store, and slots the forward did not write hold zeros
    const int i0 = std::max(1, nt_prop2 - N_taper + 1);
    const int last = std::min(nt_prop2, st.nslots);
    const int nq = last - i0 + 1;
    if (nq <= 0 || !st.d) return 0;
    const float pi = 4.0f * std::atan(1.0f);
    std::vector<float> w(nq);
    for (int n = 0; n < nq; ++n) w[n] = 0.5f * (1.0f + std::cos(pi * (float)n / (float)N_taper));
    void *v;
    if (int rc = get_buf(c, "taper_w", w.size() * sizeof(float), &v)) return rc;
    HIP_TRY(c, hipMemcpyAsync(v, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    const int lds = psdm::interior_ld(st.ny);
    psdm::launch_taper_snaps(st.d + psdm::kIntPad, (size_t)lds * st.nx, lds, st.ny, st.nx, i0 - 1, nq, (const float *)v,
                             c->stream);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));  // the host weights go out of scope
    return 0;
}

int psdm_rtm_normalize(psdm_rtm_ctx *c, const double *d_psdm_sum, const int *d_cov_sum, int ny, int nx,
                       float *d_psdm) {
    if (!c) return PSDM_RTM_EINVAL;
    if (!d_psdm_sum || !d_cov_sum || !d_psdm || ny < 1 || nx < 1) return fail(c, PSDM_RTM_EINVAL, "bad arguments");
    if (int rc = set_device(c)) return rc;
    psdm::launch_normalize(d_psdm_sum, d_cov_sum, d_psdm, (size_t)ny * nx, c->stream);
    HIP_TRY(c, hipGetLastError());
    return 0;
}

int psdm_rtm_sync(psdm_rtm_ctx *c) {
    if (!c) return PSDM_RTM_EINVAL;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

long long psdm_rtm_device_bytes(const psdm_rtm_ctx *c) {
    if (!c) return 0;
    long long n = 0;
    for (auto &kv : c->bufs) n += (long long)kv.second.n;
    for (auto &kv : c->snaps) n += (long long)kv.second.bytes;
    for (auto &sp : c->spare) n += (long long)sp.second;
    return n;
}

int psdm_rtm_set_launch(psdm_rtm_ctx *c, int tile_x, int use_graph, int graph_steps) {
    if (!c) return PSDM_RTM_EINVAL;
    if (tile_x != 8 && tile_x != 16 && tile_x != 32)
        return fail(c, PSDM_RTM_EINVAL, "tile_x must be 8, 16 or 32");
    if (graph_steps < 1) return fail(c, PSDM_RTM_EINVAL, "graph_steps must be >= 1");
    c->tile_x = tile_x;
    c->use_graph = use_graph;
    c->graph_steps = graph_steps;
    return 0;
}

int psdm_rtm_set_stream_width(psdm_rtm_ctx *c, int stream_w) {
    if (!c) return PSDM_RTM_EINVAL;
    if (stream_w < 1 || stream_w > 256) return fail(c, PSDM_RTM_EINVAL, "stream_w must be in 1..256");
    c->stream_w = stream_w;
    return 0;
}

int psdm_rtm_set_weights(psdm_rtm_ctx *c, int weights) {
    if (!c) return PSDM_RTM_EINVAL;
    if (weights != 0 && weights != 1) return fail(c, PSDM_RTM_EINVAL, "weights must be 0 (reference) or 1 (zero-sum)");
    c->weights = weights;
    return 0;
}

int psdm_rtm_set_timing(psdm_rtm_ctx *c, int timing) {
    if (!c) return PSDM_RTM_EINVAL;
    if (int rc = harvest_events(c)) return rc;
    c->timing = timing < 0 ? 0 : timing;
    c->n_timed = 0;
    c->sum_ms = 0.0;
    c->cond_ms[0] = c->cond_ms[1] = 0.0;
    c->cond_n[0] = c->cond_n[1] = 0;
    c->twt_ms[0] = c->twt_ms[1] = 0.0;
    c->twt_n = 0;
    c->su_ms[0] = c->su_ms[1] = 0.0;
    c->su_n[0] = c->su_n[1] = 0;
    c->kir_ms[0] = c->kir_ms[1] = 0.0;
    c->kir_n[0] = c->kir_n[1] = 0;
    return 0;
}

int psdm_rtm_step_stats(psdm_rtm_ctx *c, long long *n_steps, double *mean_ms) {
    if (!c) return PSDM_RTM_EINVAL;
    if (int rc = harvest_events(c)) return rc;
    if (n_steps) *n_steps = c->n_timed;
    if (mean_ms) *mean_ms = c->n_timed ? c->sum_ms / (double)c->n_timed : 0.0;
    return 0;
}

}  // extern "C"
