// Batched uneven grids: many independent problems swept by one launch per
// sweep, or per `fuse` sweeps (temporal blocking), for each group of problems
// that share a tile shape, MFMA fragment, sweeps per launch and grid size.  The
// batch owns the device-resident problem descriptors and tile lists; the caller
// owns the grids' storage buffers.
#include <stdlib.h>
#include <string.h>

#include <new>
#include <vector>

#include "tcs_internal.hpp"

using namespace tcs;

struct tcs_batch {
    int precision = 0;
    int nprob = 0;
    int64_t interior = 0;
    BatchProblem *d_probs = nullptr;
    struct Group {
        int th, tw, frag, fuse, wg_per_cu, ntiles;
        BatchTile *d_tiles;
    };
    std::vector<Group> groups;
};

static int frag_of(const tcs_problem &p) { return p.frag == 32 ? 32 : 16; }

static bool batch_tile_offered(int th, int tw, int frag) {
    if (frag == 32) return (th == 4 && tw == 4) || (th == 4 && tw == 8);
    return (th == 2 && tw == 4) || (th == 4 && tw == 4) || (th == 4 && tw == 8);
}

extern "C" int tcs_batch_destroy(tcs_batch *b) {
    if (!b) return TCS_OK;
    for (auto &g : b->groups) (void)hipFree(g.d_tiles);
    (void)hipFree(b->d_probs);
    delete b;
    return TCS_OK;
}

extern "C" int tcs_batch_create(tcs_batch **out, int nprob, const tcs_problem *probs, void *const *d_buf0,
                                void *const *d_buf1) {
    if (!out) return set_error(TCS_ERR_INVALID, "batch: null output");
    *out = nullptr;
    if (nprob < 1 || !probs || !d_buf0 || !d_buf1) return set_error(TCS_ERR_INVALID, "batch: empty or null input");
    if (probs[0].precision != TCS_PREC_FP16 && probs[0].precision != TCS_PREC_BF16)
        return set_error(TCS_ERR_UNSUPPORTED, "batch: FP16 or BF16 storage only");
    const int prec = kernel_prec(&probs[0]);  // with hi+lo coefficients or not
    std::vector<BatchProblem> hp(nprob);
    std::vector<std::vector<BatchTile>> tiles;
    struct Shape {
        int th, tw, frag, fuse, wg;
        bool operator==(const Shape &o) const {
            return th == o.th && tw == o.tw && frag == o.frag && fuse == o.fuse && wg == o.wg;
        }
    };
    std::vector<Shape> shapes;
    int64_t interior = 0;
    for (int i = 0; i < nprob; ++i) {
        const tcs_problem &p = probs[i];
        if (p.method != TCS_METHOD_MFMA || kernel_prec(&p) != prec)
            return set_error(TCS_ERR_INVALID,
                             "batch: problem %d is not MFMA with the batch's precision and coefficient form", i);
        if (has_coef_lo(prec) && frag_of(p) != 16)
            return set_error(TCS_ERR_UNSUPPORTED, "batch: hi+lo coefficients need 16x16 fragments (problem %d)", i);
        const int frag = frag_of(p);
        // 8x8 problems run on 4x8 tiles, which divide the same layout (brick
        // results depend on neither); the batched kernels stop at 4x8.
        const int th = p.tile_h == 8 && p.tile_w == 8 ? 4 : p.tile_h, tw = p.tile_w;
        if (!batch_tile_offered(th, tw, frag))
            return set_error(TCS_ERR_UNSUPPORTED,
                             "batch: problem %d tile %dx%d fragment %d (offered 2x4, 4x4, 4x8, 8x8 as 4x8; "
                             "32x32: 4x4, 4x8, 8x8 as 4x8)", i, p.tile_h, p.tile_w, frag);
        const int fuse = p.fuse > 1 ? p.fuse : 1;
        if (p.radius * (fuse - 1) > 16)
            return set_error(TCS_ERR_INVALID, "batch: problem %d fuses %d sweeps, radius %d allows 1..%d", i, fuse,
                             p.radius, 1 + 16 / p.radius);
        if (p.wg_per_cu < 0) return set_error(TCS_ERR_INVALID, "batch: problem %d has wg_per_cu < 0", i);
        const tcs_layout &L = p.layout;
        if (L.nbr % th || L.nbc % tw)
            return set_error(TCS_ERR_INVALID, "batch: problem %d layout is not a multiple of its tile", i);
        if (!d_buf0[i] || !d_buf1[i] || d_buf0[i] == d_buf1[i])
            return set_error(TCS_ERR_INVALID, "batch: problem %d needs two distinct buffers", i);
        BatchProblem &q = hp[i];
        memset(&q, 0, sizeof q);
        q.buf[0] = static_cast<uint16_t *>(d_buf0[i]);
        q.buf[1] = static_cast<uint16_t *>(d_buf1[i]);
        q.plane = L.plane_elems;
        q.rows = L.rows;
        q.cols = L.cols;
        q.sbr = L.sbr;
        q.sbc = L.sbc;
        q.rs = p.radius;
        q.tiles_x = L.nbc / tw;
        q.out_scale = ldexpf(1.0f, -p.coeff_exp);
        build_coef_windows(&p, &q.cw);
        const Shape shape{th, tw, frag, fuse, p.wg_per_cu};
        int gi = -1;
        for (size_t s = 0; s < shapes.size(); ++s)
            if (shapes[s] == shape) gi = (int)s;
        if (gi < 0) {
            shapes.push_back(shape);
            tiles.emplace_back();
            gi = (int)shapes.size() - 1;
        }
        const int nt = (L.nbr / th) * q.tiles_x;
        for (int t = 0; t < nt; ++t) tiles[gi].push_back(BatchTile{i, t});
        const int r = p.radius;
        if (L.rows > 2 * r && L.cols > 2 * r) interior += (int64_t)(L.rows - 2 * r) * (L.cols - 2 * r);
    }
    tcs_batch *b = new (std::nothrow) tcs_batch;
    if (!b) return set_error(TCS_ERR_ALLOC, "batch: out of host memory");
    b->precision = prec;
    b->nprob = nprob;
    b->interior = interior;
    hipError_t e = hipMalloc(&b->d_probs, sizeof(BatchProblem) * nprob);
    if (e == hipSuccess) e = hipMemcpy(b->d_probs, hp.data(), sizeof(BatchProblem) * nprob, hipMemcpyHostToDevice);
    for (size_t s = 0; e == hipSuccess && s < shapes.size(); ++s) {
        tcs_batch::Group g{shapes[s].th, shapes[s].tw, shapes[s].frag, shapes[s].fuse, shapes[s].wg,
                           (int)tiles[s].size(), nullptr};
        e = hipMalloc(&g.d_tiles, sizeof(BatchTile) * tiles[s].size());
        if (e == hipSuccess) {
            b->groups.push_back(g);
            e = hipMemcpy(g.d_tiles, tiles[s].data(), sizeof(BatchTile) * tiles[s].size(), hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        tcs_batch_destroy(b);
        return hip_error(e, "batch: device set-up");
    }
    *out = b;
    return TCS_OK;
}

// All `sweeps` sweeps of one group: launches of at most g.fuse sweeps,
// balanced, with a launch count of the same parity as `sweeps`, so that the
// result lands in buf[sweeps % 2] of every problem (as tcs_run does).
static int run_group(const tcs_batch *b, const tcs_batch::Group &g, int sweeps, hipStream_t stream) {
    if (sweeps <= 0) return TCS_OK;
    int per_xcd = 0;
    if (g.wg_per_cu > 0) {
        int dev = 0, cus = 0;
        TCS_HIP_CHECK(hipGetDevice(&dev));
        TCS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        per_xcd = cus * g.wg_per_cu / 8;
        if (per_xcd < 1) per_xcd = 1;
    }
    int nl = (sweeps + g.fuse - 1) / g.fuse;
    if ((nl & 1) != (sweeps & 1)) ++nl;
    const int base = sweeps / nl, extra = sweeps % nl;
    for (int i = 0; i < nl; ++i) {
        const int k = base + (i < extra ? 1 : 0);
        const int st = launch_batch(b->precision, g.th, g.tw, g.frag, b->d_probs, g.d_tiles, g.ntiles, i & 1, k,
                                    per_xcd, stream);
        if (st) return st;
    }
    return TCS_OK;
}

extern "C" int tcs_batch_run(tcs_batch *b, int sweeps, void *stream) {
    if (!b) return set_error(TCS_ERR_INVALID, "batch: null batch");
    if (sweeps < 0) return set_error(TCS_ERR_INVALID, "batch: negative sweep count");
    for (auto &g : b->groups) {
        const int st = run_group(b, g, sweeps, (hipStream_t)stream);
        if (st) return st;
    }
    return TCS_OK;
}

// Captures every group's launch chain on its own forked stream, so the graph
// holds one independent chain per group and the groups can run concurrently.
static int capture_groups(tcs_batch *b, int sweeps, hipStream_t cs) {
    if (b->groups.size() <= 1) return tcs_batch_run(b, sweeps, cs);
    std::vector<hipStream_t> side(b->groups.size(), nullptr);
    std::vector<hipEvent_t> done(b->groups.size(), nullptr);
    hipEvent_t fork = nullptr;
    int st = TCS_OK;
    hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(fork, cs);
    for (size_t i = 0; e == hipSuccess && i < side.size(); ++i) {
        e = hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamWaitEvent(side[i], fork, 0);
        if (e == hipSuccess) {
            st = run_group(b, b->groups[i], sweeps, side[i]);
            if (st) break;
        }
        if (e == hipSuccess) e = hipEventRecord(done[i], side[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(cs, done[i], 0);
    }
    if (e != hipSuccess && !st) st = hip_error(e, "batch graph: fork/join");
    for (size_t i = 0; i < side.size(); ++i) {
        if (done[i]) (void)hipEventDestroy(done[i]);
        if (side[i]) (void)hipStreamDestroy(side[i]);
    }
    if (fork) (void)hipEventDestroy(fork);
    return st;
}

extern "C" int tcs_batch_graph_create(tcs_batch *b, int sweeps, void *stream, void **graph_exec) {
    if (!graph_exec) return set_error(TCS_ERR_INVALID, "batch graph: null output");
    *graph_exec = nullptr;
    hipStream_t cs = (hipStream_t)stream;
    bool own = false;
    if (!cs) {
        TCS_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        own = true;
    }
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        if (own) (void)hipStreamDestroy(cs);
        return hip_error(e, "hipStreamBeginCapture");
    }
    const int st = b ? capture_groups(b, sweeps, cs) : set_error(TCS_ERR_INVALID, "batch: null batch");
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    if (own) (void)hipStreamDestroy(cs);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
    *graph_exec = ge;
    return TCS_OK;
}

extern "C" int64_t tcs_batch_interior_points(const tcs_batch *b) { return b ? b->interior : 0; }

extern "C" int tcs_batch_launches_per_sweep(const tcs_batch *b) { return b ? (int)b->groups.size() : 0; }
