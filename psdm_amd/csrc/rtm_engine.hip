// Batched / served RTM engine.  S shot slots stay resident on the device and advance
// together: one launch per time step walks the slot x tile work list (rtm_batch_kernel),
// and the time steps are captured in a hipGraph of `chunk_steps` launches that is
// replayed.  Between replays the host retires finished shots (image copied out and
// stacked in FP64) and admits queued shots into free slots when their snapshot store
// fits the HBM budget.  Each slot runs the same per-shot algorithm as the one-shot
// path; slots never read each other's memory, so a shot's result does not depend on
// what shares the device with it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "rtm_core.hpp"
#include "rtm_ctx.hpp"
#include "rtm_tb_host.hpp"

using namespace psdm;
using namespace psdm::host;

namespace {

// First-fit allocator over one device arena (the snapshot stores).
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_, used_;  // offset -> length
    void init(size_t n) {
        size = n;
        free_.clear();
        used_.clear();
        if (n) free_[0] = n;
    }
    bool alloc(size_t n, size_t *off) {
        n = (n + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                *off = it->first;
                const size_t rem = it->second - n, o2 = it->first + n;
                free_.erase(it);
                if (rem) free_[o2] = rem;
                used_[*off] = n;
                return true;
            }
        return false;
    }
    void release(size_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        const size_t n = u->second;
        used_.erase(u);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t in_use() const {
        size_t t = 0;
        for (auto &kv : used_) t += kv.second;
        return t;
    }
};

struct Job {
    long long id = 0;
    psdm_rtm_job j{};
    std::vector<int> nxRec, nyRec;
    size_t snap_bytes = 0;
};

struct Slot {
    bool busy = false;
    Job job;
    long long g0 = 0, g_end = 0;
    size_t arena_off = 0;
    int *d_ints = nullptr;
    size_t n_ints = 0;
    float *d_inj = nullptr;
    size_t n_inj = 0;
};

constexpr int kMaxGroups = 4;  // concurrent launch chains of one captured chunk

struct DoneRec {
    long long id;
    hipEvent_t ev;  // recorded after the shot's image is stacked
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct psdm_rtm_engine {
    psdm_rtm_ctx *c = nullptr;
    psdm_rtm_engine_cfg cfg{};
    int dn = 0, ld = 0, lds = 0, off = 0, sw = 16, nyt = 0, nxt_max = 0, cap_tiles = 0, nwork = 0, per_xcd = 0, grid = 0;
    int spl = 1;        // time steps per launch (2: temporal blocking, rtm_tblock.hip)
    int cap_core = 0;   // two steps per launch: core workgroups of a slot at capacity (cap_tiles: edge tiles)
    size_t win = 0;  // per-slot length of the image / comp buffers
    // slots split into `groups` independent launch chains inside a captured chunk, so one
    // group's next step can start while the other's drains
    int groups = 1;
    hipStream_t side[kMaxGroups - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kMaxGroups - 1] = {};
    size_t n_p = 0, n_x = 0, n_y = 0, state_base = 0, state_stride = 0, esize = 4;
    // device
    SlotDesc *d_slots = nullptr;
    long long *d_gbase = nullptr;
    float *d_tabs = nullptr, *d_state = nullptr, *d_c2 = nullptr, *d_img = nullptr, *d_comp = nullptr;
    char *d_arena = nullptr;
    double *d_sum = nullptr;
    int *d_fold = nullptr;
    bool own_image = true;
    // host
    SlotDesc *h_slots = nullptr;  // pinned, two copies (ring) for asynchronous upload
    int h_ring = 0;
    hipEvent_t h_ring_ev[2] = {nullptr, nullptr};
    bool slots_dirty = true;
    std::vector<Slot> slots;
    std::deque<Job> pending;
    Arena arena;
    long long G = 0, next_id = 1, chunks = 0;
    // one graph per work mapping (BatchArgs compact = 0, 1); compact: the mapping of the
    // uploaded slots (0 while every slot holds a full-capacity window)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int compact = 1;
    static constexpr int kRing = 256;
    hipEvent_t ev_chunk[kRing] = {};   // recorded after each chunk
    hipEvent_t ev_cstart[kRing] = {};  // recorded before each chunk (device time of the chunks)
    long long timed_upto = 0;           // chunks whose device time is in chunk_ms_sum
    hipEvent_t ev_origin = nullptr;
    double t_origin = 0.0;
    std::vector<DoneRec> done_wait;  // retired, waiting for their stack to complete
    std::vector<hipEvent_t> ev_pool;  // completion events ready for reuse
    std::vector<std::pair<long long, double>> done_ready;
    double chunk_ms_sum = 0.0;
    long long chunk_ms_n = 0;
    long long active_slot_steps = 0;  // sum over issued steps of busy slots
};

namespace {

int efail(psdm_rtm_engine *e, int code, const std::string &m) { return fail(e->c, code, m); }

// Launches one chunk of time steps (graph replay or eager launches).
// Launch arguments of slot group q (slots [q*n0, min((q+1)*n0, nslots)), n0 = ceil(nslots/groups)).
// part 1: the core kernel of two steps per launch (its own items per slot at capacity)
BatchArgs group_args(const psdm_rtm_engine *e, int q, int goff, int compact, int *grid, int part = 0) {
    const int n0 = cdiv(e->cfg.nslots, e->groups);
    const int s0 = q * n0, n = std::min(n0, e->cfg.nslots - s0);
    const int cap = part == 1 ? e->cap_core : e->cap_tiles;
    const int nwork = n * cap, per_xcd = cdiv(nwork, 8);
    *grid = e->cfg.grid > 0 ? std::min(e->cfg.grid, per_xcd * 8) : per_xcd * 8;
    return BatchArgs{e->d_slots + s0, n, e->d_gbase, goff, compact, cap, make_fastdiv(cap)};
}

// One launch of slot group q at chunk offset g: one time step, or two with temporal
// blocking (the edge-tile kernel, then the core kernel, on the same stream).
void launch_group(const psdm_rtm_engine *e, int q, int g, int compact, hipStream_t st) {
    const int sv = snap_variant(e->cfg.snap_format, e->cfg.kahan) + (e->cfg.weights == 1 ? kSvZeroSum : 0);
    int grid;
    const BatchArgs b = group_args(e, q, g, compact, &grid);
    if (e->spl == 2) {
        launch_tb_batch(b, 0, sv, grid, st);
        int grid2;
        const BatchArgs b2 = group_args(e, q, g, compact, &grid2, 1);
        launch_tb_batch(b2, 1, sv, grid2, st);
    } else {
        launch_batch(b, e->cfg.tile_x, sv, grid, st);
    }
}

// Launches a slot occupies: one per time step of each phase, or, with two steps per
// launch, nt/2 per phase (the phase's nt-1 steps two at a time).
long long slot_launches(const psdm_rtm_engine *e, int nt) { return e->spl == 2 ? 2LL * (nt / 2) : 2LL * (nt - 1); }

// Time steps that launches [l0, l1) of a slot (relative to its first) advance.
long long slot_steps_in(const psdm_rtm_engine *e, int nt, long long l0, long long l1) {
    if (l1 <= l0) return 0;
    if (e->spl != 2) return l1 - l0;
    long long n = 2 * (l1 - l0);
    if ((nt - 1) & 1) {  // the last launch of each phase advances one step
        const long long M = nt / 2;
        if (l0 <= M - 1 && M - 1 < l1) n--;
        if (l0 <= 2 * M - 1 && 2 * M - 1 < l1) n--;
    }
    return n;
}

int issue_chunk(psdm_rtm_engine *e) {
    psdm_rtm_ctx *c = e->c;
    const int C = e->cfg.chunk_steps;
    if (e->exec[e->compact]) {
        HIP_TRY(c, hipGraphLaunch(e->exec[e->compact], c->stream));
    } else {
        for (int g = 0; g < C; g++)
            for (int q = 0; q < e->groups; q++) launch_group(e, q, g, e->compact, c->stream);
        launch_advance(e->d_gbase, C, c->stream);
    }
    HIP_TRY(c, hipGetLastError());
    return 0;
}

int capture_graph(psdm_rtm_engine *e, int compact) {
    psdm_rtm_ctx *c = e->c;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    if (e->groups > 1) {  // the other chains fork off the capture stream and join before the advance
        HIP_TRY(c, hipEventRecord(e->ev_fork, c->stream));
        for (int q = 1; q < e->groups; q++) HIP_TRY(c, hipStreamWaitEvent(e->side[q - 1], e->ev_fork, 0));
    }
    for (int g = 0; g < e->cfg.chunk_steps; g++)
        for (int q = 0; q < e->groups; q++) launch_group(e, q, g, compact, q == 0 ? c->stream : e->side[q - 1]);
    for (int q = 1; q < e->groups; q++) {
        HIP_TRY(c, hipEventRecord(e->ev_join[q - 1], e->side[q - 1]));
        HIP_TRY(c, hipStreamWaitEvent(c->stream, e->ev_join[q - 1], 0));
    }
    launch_advance(e->d_gbase, e->cfg.chunk_steps, c->stream);
    HIP_TRY(c, hipStreamEndCapture(c->stream, &e->graph[compact]));
    HIP_TRY(c, hipGraphInstantiate(&e->exec[compact], e->graph[compact], nullptr, nullptr, 0));
    return 0;
}

// Snapshot store of one shot: nsnap snapshots of lds x nx_w (plus the front padding and
// a read margin), then for 16-bit formats one scale per (snapshot, stream row tile, column).
size_t snap_data_bytes(const psdm_rtm_engine *e, int nx_w, int nsnap) {
    return (((size_t)nsnap * e->lds * nx_w + 64) * e->esize + 255) & ~(size_t)255;
}
// A shot's snapshot footprint in the arena, in the arena's 256-byte granules (what
// Arena::alloc takes for it), so that a budget of S footprints holds S shots.
size_t snap_bytes_of(const psdm_rtm_engine *e, int nx_w, int nt) {
    const int nsnap = (nt - 1) / e->cfg.store_snap;
    size_t b = snap_data_bytes(e, nx_w, nsnap);
    if (e->cfg.snap_format != kSnapF32) b += (size_t)nsnap * cdiv(e->cfg.ny, kStreamRows) * nx_w * sizeof(float);
    return (b + 255) & ~(size_t)255;
}

// Builds one slot's device view for `job` starting at global step G.
int admit(psdm_rtm_engine *e, int s, Job &job, size_t arena_off) {
    psdm_rtm_ctx *c = e->c;
    const psdm_rtm_job &J = job.j;
    const int ny = e->cfg.ny, nx = J.nx_w, nt = J.nt, dn = e->dn, TX = e->cfg.tile_x;
    const int nxt = nx + 2 * dn;
    const Tiling T = make_tiling(ny, nx, dn, TX, e->sw, e->cfg.variant == 0, e->cfg.fs != 0);
    Slot &sl = e->slots[s];

    // injection events: forward source (1 point), backward receivers (last writer wins)
    auto build = [&](const int *px, const int *py, int n, std::vector<int> &ptr, std::vector<int> &cell,
                     std::vector<int> &inj, std::vector<int> &cols) {
        std::map<int, int> last;  // (iw << 16 | jw) -> trace column; ordered by column, then row
        for (int q = 0; q < n; q++) {
            const int jw = py[q] - 1, iw = px[q] - 1;
            if (jw < 0 || jw >= ny || iw < 0 || iw >= nx) continue;
            last[(iw << 16) | jw] = q;
        }
        std::vector<std::vector<std::pair<int, int>>> per(T.nstream);
        for (auto &kv : last) per[stream_item_of_cell(T, kv.first & 0xffff, kv.first >> 16)].push_back(kv);
        ptr.assign(T.nstream + 1, 0);
        for (int u = 0; u < T.nstream; u++) {
            for (auto &kv : per[u]) {
                cell.push_back(kv.first);
                inj.push_back((int)cols.size());
                cols.push_back(kv.second);
            }
            ptr[u + 1] = (int)cell.size();
        }
    };
    // two steps per launch: per work item (edge tiles, then core items) every injection cell
    // the item computes, halo included; a cell several items compute has one value column
    const TbTiling TB = make_tb_tiling(ny, nx, dn, e->sw);
    auto build_tb = [&](const int *px, const int *py, int n, std::vector<int> &ptr, std::vector<int> &cell,
                        std::vector<int> &inj, std::vector<int> &cols) {
        tb_events(TB, ny, nx, dn, px, py, n, ptr, cell, inj, cols);
    };
    const int sx = J.nxSou, sy = J.nySou;
    std::vector<int> fp, fc, fi, fcol, bp, bc, bi, bcol;
    if (e->spl == 2) {
        build_tb(&sx, &sy, 1, fp, fc, fi, fcol);
        build_tb(job.nxRec.data(), job.nyRec.data(), J.NumRec, bp, bc, bi, bcol);
    } else {
        build(&sx, &sy, 1, fp, fc, fi, fcol);
        build(job.nxRec.data(), job.nyRec.data(), J.NumRec, bp, bc, bi, bcol);
    }
    std::vector<int> ints;
    auto put = [&](const std::vector<int> &v) {
        const size_t o = ints.size();
        ints.insert(ints.end(), v.begin(), v.end());
        return o;
    };
    const size_t ofp = put(fp), ofc = put(fc), ofi = put(fi), ofcol = put(fcol);
    const size_t obp = put(bp), obc = put(bc), obi = put(bi), obcol = put(bcol);
    if (sl.n_ints < ints.size()) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (sl.d_ints) HIP_TRY(c, hipFree(sl.d_ints));
        HIP_TRY(c, hipMalloc(&sl.d_ints, ints.size() * sizeof(int)));
        sl.n_ints = ints.size();
    }
    HIP_TRY(c, hipMemcpyAsync(sl.d_ints, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const int ninj_f = (int)fcol.size(), ninj_b = (int)bcol.size();
    const size_t n_inj = (size_t)(nt - 1) * (ninj_f + ninj_b);
    if (sl.n_inj < n_inj) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (sl.d_inj) HIP_TRY(c, hipFree(sl.d_inj));
        HIP_TRY(c, hipMalloc(&sl.d_inj, std::max<size_t>(n_inj, 1) * sizeof(float)));
        sl.n_inj = std::max<size_t>(n_inj, 1);
    }
    float *inj_f = sl.d_inj, *inj_b = sl.d_inj + (size_t)(nt - 1) * ninj_f;
    launch_gather(J.d_wavelet, nt, sl.d_ints + ofcol, ninj_f, inj_f, 0, c->stream);
    launch_gather(J.d_data, nt, sl.d_ints + obcol, ninj_b, inj_b, 1, c->stream);
    // the shot's c2 window is a view of the engine's one c2 array over the whole model (the
    // reference cuts each shot's window from one Vpmodel, calculate_PSDM_RTM.f90:400-402), so
    // resident shots that overlap share their c2 cache lines
    const float *c2 = e->d_c2 + (size_t)(J.x_i - 1) * e->lds + kIntPad;
    // the host vectors above are read by the queued copy: wait before they go away
    HIP_TRY(c, hipStreamSynchronize(c->stream));

    SlotDesc &d = e->h_slots[e->h_ring * e->cfg.nslots + s];
    std::memset(&d, 0, sizeof(d));
    StepArgs &a = d.a;
    a.ny = ny;
    a.nx = nx;
    a.dn = dn;
    a.nyt = e->nyt;
    a.nxt = nxt;
    a.ld = e->ld;
    set_tiling(a, T);
    a.per_xcd = 0;
    a.fs = e->cfg.fs != 0;
    a.ram = e->cfg.variant == 0;
    a.snap_fmt = e->cfg.snap_format;
    a.kahan = e->cfg.kahan;
    a.zsum = e->cfg.weights == 1;
    a.dt = e->cfg.dt;
    a.ss_f = (float)e->cfg.store_snap;
    a.c2 = c2;
    a.ldc = e->lds;
    a.lds = e->lds;
    a.fA = e->d_tabs;
    a.fB = e->d_tabs + dn;
    a.snaps = e->d_arena + arena_off + kIntPad * e->esize;
    const int nsnap = (nt - 1) / e->cfg.store_snap;
    a.snap_scale = e->cfg.snap_format != kSnapF32 ? (float *)(e->d_arena + arena_off + snap_data_bytes(e, nx, nsnap))
                                                  : nullptr;
    a.scale_stride = T.ri * nx;
    a.image = e->d_img + (size_t)s * e->win + kIntPad;
    a.image_comp = e->d_comp + (size_t)s * e->win + kIntPad;
    a.snap_stride = (size_t)e->lds * nx;
    a.rec_row = -1;
    float *st = e->d_state + (size_t)s * e->state_stride;
    // both wavefield buffers start at zero: the _ram top/bottom pad tiles and, with a free
    // surface, the top strip are not in the work list (make_tiling), so nothing writes
    // their zeros (stream order: the
    // slot's previous shot has finished; admission is never inside a graph capture)
    HIP_TRY(c, hipMemsetAsync(st, 0, 2 * e->n_p * sizeof(float), c->stream));
    d.pA = st + e->off;
    d.pB = st + e->n_p + e->off;
    d.phxA = st + 2 * e->n_p;
    d.phxB = d.phxA + e->n_x;
    a.chx = d.phxB + e->n_x;
    d.phyA = a.chx + e->n_x;
    d.phyB = d.phyA + e->n_y;
    a.chy = d.phyB + e->n_y;
    d.hsA = a.chy + e->n_y;
    d.hsB = d.hsA + 2 * dn;
    d.g0 = e->G;
    d.nt = nt;
    d.ss = e->cfg.store_snap;
    d.fd_ss = make_fastdiv(e->cfg.store_snap);
    d.nsnap = nsnap;
    d.ev_ptr_f = sl.d_ints + ofp;
    d.ev_cell_f = sl.d_ints + ofc;
    d.ev_inj_f = sl.d_ints + ofi;
    d.inj_f = inj_f;
    d.ninj_f = ninj_f;
    d.ev_ptr_b = sl.d_ints + obp;
    d.ev_cell_b = sl.d_ints + obc;
    d.ev_inj_b = sl.d_ints + obi;
    d.inj_b = inj_b;
    d.ninj_b = ninj_b;
    if (e->spl == 2) {
        d.tb = TB;
        a.nitems = TB.ne;  // the slot's edge tiles in the launch group's first list
        d.tev_ptr_f = d.ev_ptr_f;
        d.tev_cell_f = d.ev_cell_f;
        d.tev_inj_f = d.ev_inj_f;
        d.tev_ptr_b = d.ev_ptr_b;
        d.tev_cell_b = d.ev_cell_b;
        d.tev_inj_b = d.ev_inj_b;
        float *st2 = st + e->state_base;  // the second set follows the one-step state
        d.pC = st2 + e->off;
        d.pD = st2 + e->n_p + e->off;
        float *x2 = st2 + 2 * e->n_p;
        d.chxB = x2;
        d.chyB = x2 + e->n_x;
        d.hsC = d.chyB + e->n_y;
        d.hsD = d.hsC + 2 * dn;
    }
    sl.busy = true;
    sl.g0 = e->G;
    sl.g_end = e->G + slot_launches(e, nt);
    sl.arena_off = arena_off;
    sl.job = std::move(job);
    e->slots_dirty = true;
    return 0;
}

int retire(psdm_rtm_engine *e, int s) {
    psdm_rtm_ctx *c = e->c;
    Slot &sl = e->slots[s];
    const psdm_rtm_job &J = sl.job.j;
    const int ny = e->cfg.ny;
    float *img = e->d_img + (size_t)s * e->win + kIntPad;
    float *comp = e->d_comp + (size_t)s * e->win + kIntPad;
    if (J.d_image_out)
        HIP_TRY(c, hipMemcpy2DAsync(J.d_image_out, (size_t)ny * sizeof(float), img, (size_t)e->lds * sizeof(float),
                                    (size_t)ny * sizeof(float), J.nx_w, hipMemcpyDeviceToDevice, c->stream));
    launch_stack(img, e->cfg.kahan ? comp : nullptr, e->lds, ny, J.nx_w, J.x_i - 1, e->d_sum, ny, e->d_fold,
                 c->stream);
    e->arena.release(sl.arena_off);
    hipEvent_t ev;
    if (!e->ev_pool.empty()) {
        ev = e->ev_pool.back();
        e->ev_pool.pop_back();
    } else {
        HIP_TRY(c, hipEventCreate(&ev));
    }
    HIP_TRY(c, hipEventRecord(ev, c->stream));
    e->done_wait.push_back({sl.job.id, ev});
    sl.busy = false;
    SlotDesc &d = e->h_slots[e->h_ring * e->cfg.nslots + s];
    d.g0 = -1;
    e->slots_dirty = true;
    return 0;
}

int upload_slots(psdm_rtm_engine *e) {
    psdm_rtm_ctx *c = e->c;
    if (!e->slots_dirty) return 0;
    const int S = e->cfg.nslots;
    SlotDesc *src = e->h_slots + e->h_ring * S;
    const int n0 = cdiv(S, e->groups);  // compact work list of each launch group (group_args)
    e->compact = 0;
    for (int s0 = 0; s0 < S; s0 += n0) {
        int run = 0;
        const int n = std::min(S, s0 + n0) - s0;
        bool uni = true;
        for (int s = s0; s < s0 + n; ++s) {
            src[s].woff = run;
            src[s].witems = src[s].g0 >= 0 ? src[s].a.nitems : 0;
            run += src[s].witems;
            uni = uni && src[s].witems == src[s0].witems;
        }
        src[s0].wtotal = run;
        src[s0].wuni = uni ? src[s0].witems : 0;
        src[s0].wfd = make_fastdiv(std::max(1, run / n));
        if (src[s0].wuni != e->cap_tiles) e->compact = 1;
        if (e->spl == 2) {  // the core kernel's list
            int run2 = 0, first = -1;
            bool uni2 = true;
            for (int s = s0; s < s0 + n; ++s) {
                const int w2 = src[s].g0 >= 0 ? src[s].tb.ncblk : 0;
                src[s].cwoff = run2;
                run2 += w2;
                if (first < 0) first = w2;
                uni2 = uni2 && w2 == first;
            }
            src[s0].cwtotal = run2;
            src[s0].cwuni = uni2 ? first : 0;
            src[s0].cwfd = make_fastdiv(std::max(1, run2 / n));
            if (src[s0].cwuni != e->cap_core) e->compact = 1;
        }
    }
    HIP_TRY(c, hipMemcpyAsync(e->d_slots, src, S * sizeof(SlotDesc), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipEventRecord(e->h_ring_ev[e->h_ring], c->stream));
    // continue editing in the other copy once its previous upload has been consumed
    const int nxt = e->h_ring ^ 1;
    HIP_TRY(c, hipEventSynchronize(e->h_ring_ev[nxt]));
    std::memcpy(e->h_slots + nxt * S, src, S * sizeof(SlotDesc));
    e->h_ring = nxt;
    e->slots_dirty = false;
    return 0;
}

// Reports jobs whose last chunk has completed.  Entries whose chunk is at or before
// `block_upto` are waited for (their ring event is about to be reused); the rest are
// only queried.
// Folds the device time of completed chunks (start/stop events around each graph of
// chunk_steps launches) into chunk_ms_sum, waiting for the chunks up to `block_upto`.
int harvest_chunk_times(psdm_rtm_engine *e, long long block_upto) {
    psdm_rtm_ctx *c = e->c;
    while (e->timed_upto < e->chunks) {
        const long long q = e->timed_upto;
        hipEvent_t e0 = e->ev_cstart[q % psdm_rtm_engine::kRing], e1 = e->ev_chunk[q % psdm_rtm_engine::kRing];
        const hipError_t st = q <= block_upto ? hipEventSynchronize(e1) : hipEventQuery(e1);
        if (st == hipErrorNotReady) {
            (void)hipGetLastError();  // a pending event is not an error of the caller's
            break;
        }
        if (st != hipSuccess) return hip_fail(c, st, "chunk event");
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
        e->chunk_ms_sum += ms;
        e->chunk_ms_n++;
        e->timed_upto++;
    }
    return 0;
}

// Moves retired shots whose image stack has completed on the device to done_ready
// (waiting for them when `block`), with the completion time on the host steady clock.
int collect(psdm_rtm_engine *e, bool block) {
    psdm_rtm_ctx *c = e->c;
    std::vector<DoneRec> keep;
    for (auto &d : e->done_wait) {
        hipError_t q = block ? hipEventSynchronize(d.ev) : hipEventQuery(d.ev);
        if (q == hipSuccess) {
            float ms = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&ms, e->ev_origin, d.ev));
            e->done_ready.emplace_back(d.id, e->t_origin + 1e-3 * ms);
            e->ev_pool.push_back(d.ev);
        } else if (q == hipErrorNotReady) {
            (void)hipGetLastError();
            keep.push_back(d);
        } else {
            return hip_fail(c, q, "shot completion event");
        }
    }
    e->done_wait.swap(keep);
    return 0;
}

}  // namespace

extern "C" {

int psdm_rtm_engine_create(psdm_rtm_ctx *c, const psdm_rtm_engine_cfg *cfg, psdm_rtm_engine **out) {
    if (!c || !cfg || !out) return PSDM_RTM_EINVAL;
    *out = nullptr;
    if (int rc = set_device(c)) return rc;
    if (cfg->nslots < 1 || cfg->nslots > 256) return fail(c, PSDM_RTM_EINVAL, "nslots must be in 1..256");
    if (cfg->ny < 1 || cfg->max_nx < 1 || cfg->max_nt < 2 || cfg->nx_prop < 1)
        return fail(c, PSDM_RTM_EINVAL, "bad engine capacity");
    if (cfg->ny >= 65536 || cfg->max_nx >= 32768) return fail(c, PSDM_RTM_EINVAL, "window exceeds 65535 x 32767 cells");
    if (cfg->dPML < 2 || cfg->dPML + 6 > kMaxTables) return fail(c, PSDM_RTM_EINVAL, "dPML out of range");
    if (cfg->store_snap < 1) return fail(c, PSDM_RTM_EINVAL, "store_snap must be >= 1");
    if (cfg->snap_format < 0 || cfg->snap_format > 2) return fail(c, PSDM_RTM_EINVAL, "snap_format must be 0, 1 or 2");
    if (cfg->tile_x != 8 && cfg->tile_x != 16 && cfg->tile_x != 32)
        return fail(c, PSDM_RTM_EINVAL, "tile_x must be 8, 16 or 32");
    if (cfg->stream_w < 0 || cfg->stream_w > 256) return fail(c, PSDM_RTM_EINVAL, "stream_w must be in 0..256");
    if (cfg->groups < 0 || cfg->groups > kMaxGroups) return fail(c, PSDM_RTM_EINVAL, "groups must be in 0..4");
    if (cfg->chunk_steps < 1 || cfg->chunk_steps > 4096) return fail(c, PSDM_RTM_EINVAL, "chunk_steps out of range");
    if (cfg->steps_per_launch < 0 || cfg->steps_per_launch > 2)
        return fail(c, PSDM_RTM_EINVAL, "steps_per_launch must be 0, 1 or 2");
    if (cfg->steps_per_launch == 2 && cfg->snap_format != kSnapF32)
        return fail(c, PSDM_RTM_EINVAL, "steps_per_launch 2 needs FP32 snapshots");
    if (cfg->steps_per_launch == 2 && cfg->ny < 16)
        return fail(c, PSDM_RTM_EINVAL, "steps_per_launch 2 needs ny >= 16");
    if (cfg->steps_per_launch == 2 && cfg->dPML + 6 > 256)
        return fail(c, PSDM_RTM_EINVAL, "steps_per_launch 2 needs dPML <= 250");
    if (cfg->weights != 0 && cfg->weights != 1)
        return fail(c, PSDM_RTM_EINVAL, "weights must be 0 (reference) or 1 (zero-sum)");
    if (cfg->weights == 1 && cfg->steps_per_launch == 2)
        return fail(c, PSDM_RTM_EINVAL, "zero-sum weights need steps_per_launch <= 1");
    if (!cfg->d_vel) return fail(c, PSDM_RTM_EINVAL, "null velocity model");
    if (!(cfg->dt > 0.f) || !(cfg->dmodel > 0.f)) return fail(c, PSDM_RTM_EINVAL, "dt and dmodel must be > 0");
    auto *e = new psdm_rtm_engine();
    e->c = c;
    e->cfg = *cfg;
    e->dn = cfg->dPML + 6;
    e->nyt = cfg->ny + 2 * e->dn;
    e->ld = state_ld(cfg->ny, e->dn);
    e->spl = cfg->steps_per_launch == 2 ? 2 : 1;
    // two steps per launch: core items read 4 rows past the one-step items' last (interior
    // rows up to ny + 7), so the interior arrays hold 4 more rows
    e->lds = e->spl == 2 ? cdiv((cfg->ny > kStreamRows ? cfg->ny : kStreamRows) + 12, 64) * 64
                         : psdm::interior_ld(cfg->ny);
    e->off = state_off(e->dn);
    e->sw = cfg->stream_w > 0 ? cfg->stream_w : (e->spl == 2 ? 8 : 4);
    e->nxt_max = cfg->max_nx + 2 * e->dn;
    if (e->spl == 2) {
        const TbTiling T = make_tb_tiling(cfg->ny, cfg->max_nx, e->dn, e->sw);
        e->cap_tiles = T.ne;
        e->cap_core = T.ncblk;
    } else {
        e->cap_tiles = make_tiling(cfg->ny, cfg->max_nx, e->dn, cfg->tile_x, e->sw, cfg->variant == 0, cfg->fs != 0).nitems;
    }
    e->groups = cfg->groups > 0 ? cfg->groups : 2;
    if (!cfg->use_graph) e->groups = 1;
    e->groups = cdiv(cfg->nslots, cdiv(cfg->nslots, e->groups));  // every group holds at least one slot
    e->win = (size_t)e->lds * cfg->max_nx + 64;
    e->nwork = cfg->nslots * e->cap_tiles;
    e->per_xcd = (e->nwork + 7) / 8;
    e->grid = cfg->grid > 0 ? std::min(cfg->grid, e->per_xcd * 8) : e->per_xcd * 8;
    e->esize = cfg->snap_format == kSnapF32 ? 4 : 2;
    const size_t dn2 = 2 * e->dn;
    e->n_p = ((size_t)e->ld * e->nxt_max + 32 + 63) / 64 * 64;  // + room for the row offset
    e->n_x = (size_t)e->ld * dn2;
    e->n_y = dn2 * e->nxt_max;
    // p (2 levels), phi_x (2), chi_x, phi_y (2), chi_y, the x-strip history (2); with two
    // steps per launch a second pair of p levels, second chi buffers and history pair
    e->state_base = (2 * e->n_p + 3 * e->n_x + 3 * e->n_y + 2 * dn2 + 63) / 64 * 64;
    e->state_stride = e->state_base;
    if (e->spl == 2) e->state_stride += (2 * e->n_p + e->n_x + e->n_y + 2 * dn2 + 63) / 64 * 64;
    const size_t S = cfg->nslots, win = e->win;
    size_t budget = cfg->hbm_budget > 0 ? (size_t)cfg->hbm_budget : S * snap_bytes_of(e, cfg->max_nx, cfg->max_nt);
    auto bail = [&](int rc) {
        psdm_rtm_engine_destroy(e);
        return rc;
    };
#define ETRY(expr)                                                                                \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return bail(hip_fail(c, e_, #expr));                                                  \
    } while (0)
    if (e->groups > 1) ETRY(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    for (int q = 1; q < e->groups; q++) {
        ETRY(hipStreamCreateWithFlags(&e->side[q - 1], hipStreamNonBlocking));
        ETRY(hipEventCreateWithFlags(&e->ev_join[q - 1], hipEventDisableTiming));
    }
    ETRY(hipMalloc(&e->d_slots, S * sizeof(SlotDesc)));
    ETRY(hipMalloc(&e->d_gbase, sizeof(long long)));
    ETRY(hipMalloc(&e->d_tabs, 2 * e->dn * sizeof(float)));
    ETRY(hipMalloc(&e->d_state, S * e->state_stride * sizeof(float)));
    // c2 = v^2 dt^2 / dx^2 of the whole model, in the interior layout (lds rows per column)
    ETRY(hipMalloc(&e->d_c2, ((size_t)e->lds * cfg->nx_prop + 64) * sizeof(float)));
    launch_c2(cfg->d_vel, e->d_c2 + kIntPad, cfg->ny, cfg->nx_prop, e->lds, cfg->dt, cfg->dmodel, c->stream);
    ETRY(hipMalloc(&e->d_img, S * win * sizeof(float)));
    ETRY(hipMalloc(&e->d_comp, S * win * sizeof(float)));
    ETRY(hipMalloc(&e->d_arena, std::max<size_t>(budget, 256)));
    if (cfg->d_psdm_sum && cfg->d_cov_sum) {
        e->d_sum = cfg->d_psdm_sum;
        e->d_fold = cfg->d_cov_sum;
        e->own_image = false;
    } else {
        ETRY(hipMalloc(&e->d_sum, (size_t)cfg->ny * cfg->nx_prop * sizeof(double)));
        ETRY(hipMalloc(&e->d_fold, (size_t)cfg->ny * cfg->nx_prop * sizeof(int)));
        ETRY(hipMemsetAsync(e->d_sum, 0, (size_t)cfg->ny * cfg->nx_prop * sizeof(double), c->stream));
        ETRY(hipMemsetAsync(e->d_fold, 0, (size_t)cfg->ny * cfg->nx_prop * sizeof(int), c->stream));
    }
    ETRY(hipHostMalloc((void **)&e->h_slots, 2 * S * sizeof(SlotDesc), hipHostMallocDefault));
    ETRY(hipMemsetAsync(e->d_img, 0, S * win * sizeof(float), c->stream));
    ETRY(hipMemsetAsync(e->d_comp, 0, S * win * sizeof(float), c->stream));
    ETRY(hipMemsetAsync(e->d_gbase, 0, sizeof(long long), c->stream));
    std::vector<float> fA, fB;
    pml_profile(cfg->dPML, (double)cfg->dt, (double)cfg->dmodel, fA, fB);
    fA.insert(fA.end(), fB.begin(), fB.end());
    ETRY(hipMemcpyAsync(e->d_tabs, fA.data(), fA.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    std::memset(e->h_slots, 0, 2 * S * sizeof(SlotDesc));
    for (size_t s = 0; s < 2 * S; s++) e->h_slots[s].g0 = -1;
    for (int q = 0; q < 2; q++) ETRY(hipEventCreate(&e->h_ring_ev[q]));
    for (int q = 0; q < psdm_rtm_engine::kRing; q++) ETRY(hipEventCreate(&e->ev_chunk[q]));
    for (int q = 0; q < psdm_rtm_engine::kRing; q++) ETRY(hipEventCreate(&e->ev_cstart[q]));
    ETRY(hipEventCreate(&e->ev_origin));
    e->slots.resize(S);
    e->arena.init(budget);
    if (int rc = upload_slots(e)) return bail(rc);
    ETRY(hipStreamSynchronize(c->stream));
    if (cfg->use_graph)
        for (int m = 0; m < 2; m++)
            if (int rc = capture_graph(e, m)) return bail(rc);
    ETRY(hipEventRecord(e->ev_origin, c->stream));
    ETRY(hipEventSynchronize(e->ev_origin));
    e->t_origin = now_s();
#undef ETRY
    *out = e;
    return 0;
}

int psdm_rtm_engine_destroy(psdm_rtm_engine *e) {
    if (!e) return 0;
    psdm_rtm_ctx *c = e->c;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (int m = 0; m < 2; m++) {
        if (e->exec[m]) (void)hipGraphExecDestroy(e->exec[m]);
        if (e->graph[m]) (void)hipGraphDestroy(e->graph[m]);
    }
    for (auto &sl : e->slots) {
        if (sl.d_ints) (void)hipFree(sl.d_ints);
        if (sl.d_inj) (void)hipFree(sl.d_inj);
    }
    void *bufs[] = {e->d_slots, e->d_gbase, e->d_tabs, e->d_state, e->d_c2, e->d_img, e->d_comp, e->d_arena};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (e->own_image) {
        if (e->d_sum) (void)hipFree(e->d_sum);
        if (e->d_fold) (void)hipFree(e->d_fold);
    }
    if (e->h_slots) (void)hipHostFree(e->h_slots);
    for (auto ev : e->h_ring_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto ev : e->ev_chunk)
        if (ev) (void)hipEventDestroy(ev);
    for (auto ev : e->ev_cstart)
        if (ev) (void)hipEventDestroy(ev);
    if (e->ev_origin) (void)hipEventDestroy(e->ev_origin);
    for (auto ev : e->ev_pool) (void)hipEventDestroy(ev);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    for (int q = 0; q < kMaxGroups - 1; q++) {
        if (e->ev_join[q]) (void)hipEventDestroy(e->ev_join[q]);
        if (e->side[q]) (void)hipStreamDestroy(e->side[q]);
    }
    for (auto &d : e->done_wait) (void)hipEventDestroy(d.ev);
    delete e;
    return 0;
}

int psdm_rtm_engine_submit(psdm_rtm_engine *e, const psdm_rtm_job *job, long long *id) {
    if (!e || !job) return PSDM_RTM_EINVAL;
    psdm_rtm_ctx *c = e->c;
    if (job->nx_w < 1 || job->nx_w > e->cfg.max_nx) return efail(e, PSDM_RTM_EINVAL, "window wider than the slot capacity");
    if (job->nt < 2 || job->nt > e->cfg.max_nt) return efail(e, PSDM_RTM_EINVAL, "record length outside 2..max_nt");
    if (job->x_i < 1 || job->x_i + job->nx_w - 1 > e->cfg.nx_prop)
        return efail(e, PSDM_RTM_EINVAL, "window outside the velocity model");
    if (e->cfg.variant == 0 && job->nx_w < e->dn)
        return efail(e, PSDM_RTM_EINVAL, "the _ram padding needs nx_w >= dPML+6");
    if (e->spl == 2 && job->nx_w < 16) return efail(e, PSDM_RTM_EINVAL, "steps_per_launch 2 needs nx_w >= 16");
    if (job->NumRec < 0 || (job->NumRec > 0 && (!job->nxRec || !job->nyRec || !job->d_data)) || !job->d_wavelet)
        return efail(e, PSDM_RTM_EINVAL, "missing gather, receivers or wavelet");
    Job J;
    J.id = e->next_id++;
    J.j = *job;
    J.nxRec.assign(job->nxRec, job->nxRec + job->NumRec);
    J.nyRec.assign(job->nyRec, job->nyRec + job->NumRec);
    J.j.nxRec = nullptr;
    J.j.nyRec = nullptr;
    J.snap_bytes = snap_bytes_of(e, job->nx_w, job->nt);
    if (J.snap_bytes > e->arena.size)
        return efail(e, PSDM_RTM_EINVAL, "the shot's snapshot store exceeds the engine's HBM budget");
    if (id) *id = J.id;
    e->pending.push_back(std::move(J));
    (void)c;
    return 0;
}

int psdm_rtm_engine_run(psdm_rtm_engine *e, int max_chunks, int *n_active, int *n_pending) {
    if (!e) return PSDM_RTM_EINVAL;
    psdm_rtm_ctx *c = e->c;
    if (int rc = set_device(c)) return rc;
    const int C = e->cfg.chunk_steps;
    const int inflight = std::max(1, e->cfg.max_inflight);
    for (int n = 0; n < max_chunks; n++) {
        // retire shots whose last step is in an issued chunk
        for (int s = 0; s < e->cfg.nslots; s++)
            if (e->slots[s].busy && e->slots[s].g_end <= e->G)
                if (int rc = retire(e, s)) return rc;
        // admit queued shots into free slots, first fit against the HBM budget
        for (int s = 0; s < e->cfg.nslots && !e->pending.empty(); s++) {
            if (e->slots[s].busy) continue;
            for (auto it = e->pending.begin(); it != e->pending.end(); ++it) {
                size_t off;
                if (!e->arena.alloc(it->snap_bytes, &off)) continue;
                Job job = std::move(*it);
                e->pending.erase(it);
                if (int rc = admit(e, s, job, off)) return rc;
                break;
            }
        }
        bool any = false;
        for (auto &sl : e->slots) any = any || sl.busy;
        if (!any) {
            // retire releases a shot's snapshots at once, so the arena is empty here: a
            // queued shot that was not admitted never will be
            if (!e->pending.empty())
                return efail(e, PSDM_RTM_EINVAL, "a queued shot does not fit the empty snapshot arena");
            break;
        }
        if (int rc = upload_slots(e)) return rc;
        // bounded look-ahead: wait for the chunk issued `inflight` chunks ago
        if (e->chunks >= inflight) {
            hipEvent_t old = e->ev_chunk[(e->chunks - inflight) % psdm_rtm_engine::kRing];
            HIP_TRY(c, hipEventSynchronize(old));
        }
        // the ring entry about to be re-recorded must have its device time folded in
        if (int rc = harvest_chunk_times(e, e->chunks - psdm_rtm_engine::kRing)) return rc;
        HIP_TRY(c, hipEventRecord(e->ev_cstart[e->chunks % psdm_rtm_engine::kRing], c->stream));
        if (int rc = issue_chunk(e)) return rc;
        for (auto &sl : e->slots)
            if (sl.busy) e->active_slot_steps += slot_steps_in(e, sl.job.j.nt, e->G - sl.g0, std::min(e->G + C, sl.g_end) - sl.g0);
        e->G += C;
        HIP_TRY(c, hipEventRecord(e->ev_chunk[e->chunks % psdm_rtm_engine::kRing], c->stream));
        e->chunks++;
    }
    if (int rc = collect(e, false)) return rc;
    int na = 0;
    for (auto &sl : e->slots) na += sl.busy;
    if (n_active) *n_active = na;
    if (n_pending) *n_pending = (int)e->pending.size();
    return 0;
}

int psdm_rtm_engine_poll(psdm_rtm_engine *e, long long *ids, double *t_done, int cap, int *n) {
    if (!e) return PSDM_RTM_EINVAL;
    if (int rc = collect(e, false)) return rc;
    const int k = std::min<int>(cap, (int)e->done_ready.size());
    for (int q = 0; q < k; q++) {
        if (ids) ids[q] = e->done_ready[q].first;
        if (t_done) t_done[q] = e->done_ready[q].second;
    }
    e->done_ready.erase(e->done_ready.begin(), e->done_ready.begin() + k);
    if (n) *n = k;
    return 0;
}

int psdm_rtm_engine_drain(psdm_rtm_engine *e) {
    if (!e) return PSDM_RTM_EINVAL;
    int na = 1, np = 1;
    while (na > 0 || np > 0)
        if (int rc = psdm_rtm_engine_run(e, 16, &na, &np)) return rc;
    // retire the last shots and wait for everything
    if (int rc = psdm_rtm_engine_run(e, 1, &na, &np)) return rc;
    HIP_TRY(e->c, hipStreamSynchronize(e->c->stream));
    return collect(e, true);
}

int psdm_rtm_engine_image(psdm_rtm_engine *e, double **d_psdm_sum, int **d_cov_sum) {
    if (!e) return PSDM_RTM_EINVAL;
    if (d_psdm_sum) *d_psdm_sum = e->d_sum;
    if (d_cov_sum) *d_cov_sum = e->d_fold;
    return 0;
}

int psdm_rtm_engine_timing(psdm_rtm_engine *e, double *mean_launch_ms, long long *n_launches) {
    if (!e) return PSDM_RTM_EINVAL;
    if (int rc = harvest_chunk_times(e, e->chunks - 1)) return rc;
    const long long n = e->chunk_ms_n * e->cfg.chunk_steps;
    if (n_launches) *n_launches = n;
    if (mean_launch_ms) *mean_launch_ms = n > 0 ? e->chunk_ms_sum / (double)n : 0.0;
    return 0;
}

int psdm_rtm_engine_memory(psdm_rtm_engine *e, long long *fixed_bytes, long long *arena_bytes) {
    if (!e) return PSDM_RTM_EINVAL;
    const size_t S = e->cfg.nslots;
    size_t fixed = S * sizeof(SlotDesc) + sizeof(long long) + 2 * e->dn * sizeof(float);
    fixed += S * e->state_stride * sizeof(float) + 2 * S * e->win * sizeof(float);
    fixed += ((size_t)e->lds * e->cfg.nx_prop + 64) * sizeof(float);
    if (e->own_image) fixed += (size_t)e->cfg.ny * e->cfg.nx_prop * (sizeof(double) + sizeof(int));
    for (auto &sl : e->slots) fixed += sl.n_ints * sizeof(int) + sl.n_inj * sizeof(float);
    if (fixed_bytes) *fixed_bytes = (long long)fixed;
    if (arena_bytes) *arena_bytes = (long long)e->arena.size;
    return 0;
}

int psdm_rtm_engine_stats(psdm_rtm_engine *e, long long *steps_issued, long long *slot_steps, long long *hbm_used,
                          long long *hbm_budget) {
    if (!e) return PSDM_RTM_EINVAL;
    if (steps_issued) *steps_issued = e->G;
    if (slot_steps) *slot_steps = e->active_slot_steps;
    if (hbm_used) *hbm_used = (long long)e->arena.in_use();
    if (hbm_budget) *hbm_budget = (long long)e->arena.size;
    return 0;
}

}  // extern "C"
