// Batch plans: many molecules through grid build -> FIRE -> df as one
// captured HIP graph, plus the autotuner of the persistent relaxation.
//
// The reference evaluates one molecule per call (AFMulator.__call__,
// ppafm/ocl/AFMulator.py:142-179, one blocking kernel sequence per
// molecule; relaxTipStrokes_omp, ppafm/cpp/ProbeParticle.cpp:682-701).  A
// plan instead lays every grid and every output back to back in HBM and
// issues, per run, four kernels:
//   fill      clears the relaxation work queues and counters (a kernel, not
//             a memset node, so graph replays stay correct);
//   grid      one launch over the bricks of all grids (block -> molecule by
//             binary search over brick prefix sums);
//   relax     one persistent launch, waves pull pixel tiles from per-XCD
//             queues;
//   df        one z-correlation launch per run of molecules with equal nz.
// The four are captured once into a graph and replayed by ppafm_batch_run.

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "pp_common.hpp"

using namespace pp;

namespace {

struct DfGroup {
    int first, last;  // items [first, last)
    int nz;
    long long ncol;
    long long pos_off, df_off;
};

int device_cus(int &ncu)
{
    int dev = 0;
    PP_TRY_HIP(hipGetDevice(&dev));
    PP_TRY_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (ncu < 1) ncu = 1;
    return 0;
}

template <typename T>
int upload(T *&dst, const std::vector<T> &src, hipStream_t st)
{
    if (!dst) PP_TRY_HIP(hipMalloc((void **)&dst, std::max<size_t>(src.size(), 1) * sizeof(T)));
    if (!src.empty()) PP_TRY_HIP(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return 0;
}

}  // namespace

struct ppafm_batch {
    int nitems = 0;
    std::vector<ppafm_batch_item> items;
    ppafm_ff_model model{};
    ppafm_relax_params rp{};
    ppafm_tip tip{};
    bool f64 = true;
    std::vector<double> w;
    double df_scale = 1;
    bool keep_iters = false;
    ppafm_batch_tuning tuning{8, 1, 16, 0};
    bool homog = false;  // every item shares grid shape, cell and scan (lane refill possible)
    std::vector<BatchRelaxItem> ritems;  // host copy of the relaxation table
    hipStream_t st = nullptr;
    bool own_stream = false;
    int ncu = 1;

    std::vector<long long> grid_off, pos_off, df_off, atom_off, brick_prefix;
    std::vector<BatchGridItem> gitems;
    std::vector<DfGroup> groups;

    void *grids = nullptr;
    float *pos = nullptr;
    float4 *force = nullptr;
    int *iters = nullptr;
    unsigned char *status = nullptr;  // per-position status, with iters
    float *df = nullptr;
    AtomRec *atoms = nullptr;
    long long atom_cap = 0;  // atom records the atoms buffer holds
    unsigned *ctl = nullptr;  // 8 queue heads (u32) then 4 counters (u64)
    BatchGridItem *d_gitems = nullptr;
    long long *d_bprefix = nullptr;
    BatchRelaxItem *d_ritems = nullptr;
    long long *d_tprefix = nullptr;
    long long ntiles = 0;
    // tuning scratch tables
    BatchRelaxItem *d_sitems = nullptr;
    long long *d_sprefix = nullptr;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    bool any_uncharged = false;  // some uploaded atom has q = 0 (grid kernel variant)

    // warm starts (ppafm_batch_set_warmstart)
    ppafm_warm_params warm{};
    ppafm_ws_model *ws = nullptr;
    float *pred = nullptr;           // step-0 estimates (pixel columns, 3)
    long long *d_wprefix = nullptr;  // 8x8-column tiles of items < k (step-0 estimates)
    std::vector<long long> wprefix;  // host copy
    long long nwaves = 0;
    long long *d_pprefix = nullptr;  // positions of items < k

    unsigned long long *counters() const { return (unsigned long long *)(ctl + 8); }
};

namespace {

// Relaxation item table and tile prefix for the first n items under a tile shape.
void relax_tables(const ppafm_batch &b, int n, int tile_y, std::vector<BatchRelaxItem> &it,
                  std::vector<long long> &prefix)
{
    it.resize(n);
    prefix.assign(n + 1, 0);
    long long ncol = 0;
    for (int k = 0; k < n; k++) {
        const ppafm_batch_item &s = b.items[k];
        BatchRelaxItem &r = it[k];
        std::memset(&r, 0, sizeof(r));
        GridGeom g = make_geom(s.n, s.cell);
        for (int a = 0; a < 3; a++) r.n[a] = s.n[a];
        r.nx = s.scan.nx; r.ny = s.scan.ny; r.nz = s.scan.nz;
        const long long nt = (r.nz > 0) ? relax_batch_tiles(r.nx, r.ny, tile_y, r.tiles_y) : 0;
        if (r.nz <= 0) r.tiles_y = 1;
        r.grid_off = b.grid_off[k];
        r.out_off = b.pos_off[k];
        r.col_off = (int)ncol;
        if (r.nz > 0) ncol += (long long)r.nx * r.ny;
        for (int a = 0; a < 9; a++) r.dinv[a] = g.dinv[a];
        for (int a = 0; a < 3; a++) {
            r.p0[a] = s.scan.p0[a]; r.da[a] = s.scan.da[a]; r.db[a] = s.scan.db[a]; r.dstep[a] = s.scan.dstep[a];
        }
        for (int a = 0; a < 9; a++) r.dcell[a] = g.dcell[a];
        prefix[k + 1] = prefix[k] + nt;
    }
}

// Every item shares grid shape, cell and scan (only offsets differ): the
// relaxation can refill lanes column by column with one set of uniforms.
bool homogeneous(const std::vector<BatchRelaxItem> &it)
{
    if (it.empty()) return false;
    const BatchRelaxItem &a = it[0];
    if (a.nz <= 0) return false;
    for (const BatchRelaxItem &r : it) {
        if (r.nx != a.nx || r.ny != a.ny || r.nz != a.nz || r.tiles_y != a.tiles_y) return false;
        for (int k = 0; k < 3; k++)
            if (r.n[k] != a.n[k] || r.p0[k] != a.p0[k] || r.da[k] != a.da[k] || r.db[k] != a.db[k] || r.dstep[k] != a.dstep[k])
                return false;
        for (int k = 0; k < 9; k++)
            if (r.dinv[k] != a.dinv[k]) return false;
    }
    return true;
}

int enqueue_grid(ppafm_batch &b, int nitems, long long nbricks)
{
    bool ortho = true;
    for (const BatchGridItem &g : b.gitems) ortho = ortho && ortho_cell(g.dcell);
    return launch_ff_build_batch(b.model, nitems, b.d_gitems, b.d_bprefix, nbricks, b.atoms, b.grids, b.f64,
                                 b.any_uncharged, ortho, b.st);
}

int enqueue_relax(ppafm_batch &b, const ppafm_batch_tuning &t, const BatchRelaxItem *items, const long long *prefix,
                  int nitems, long long ntiles, const std::vector<BatchRelaxItem> &hitems)
{
    RelaxBatchLaunch L;
    L.refill = t.refill >= 1 && homogeneous(hitems);
    L.refill_chunk = L.refill && t.refill >= 2 ? 64 * (t.refill - 1) : 0;
    if (L.refill) L.item0 = hitems[0];
    // the orthogonal kernels: every item's dinv diagonal and its grid under
    // 4 GiB (they read grids through buffer resources with 32-bit offsets)
    L.ortho = true;
    for (const BatchRelaxItem &r : hitems) {
        for (int k : {1, 2, 3, 5, 6, 7}) L.ortho = L.ortho && r.dinv[k] == 0;
        L.ortho = L.ortho && (unsigned long long)r.n[0] * r.n[1] * r.n[2] * (b.f64 ? 32 : 16) <= 0xFFFFFFFFull;
    }
    L.rp = b.rp;
    L.tip = b.tip;
    L.items = items;
    L.prefix = prefix;
    L.nitems = nitems;
    L.ntiles = ntiles;
    L.grids = b.grids;
    L.f64 = b.f64;
    L.queue = b.ctl;
    L.pos = b.pos;
    L.force = b.force;
    L.iters = b.iters;
    L.status = b.status;
    L.counters = b.counters();
    L.tile_y = t.tile_y;
    L.waves_per_block = t.waves_per_block;
    L.nblocks = std::max(1, b.ncu * t.blocks_per_cu);
    if (b.warm.enable) {
        L.warm = true;
        L.pred = b.ws ? b.pred : nullptr;
        L.newton_max = b.warm.newton_max;
        L.newton_max0 = b.warm.newton_max0;
        L.newton_retry = b.warm.newton_retry;
        L.newton_tol = b.warm.newton_tol;
        L.newton_cap = b.warm.newton_cap;
        L.newton_extrap = b.warm.extrap_max;
        L.residual_max = b.warm.residual_max;
    }
    return launch_relax_batch(L, b.st);
}

RelaxBatchLaunch base_launch(ppafm_batch &b)
{
    RelaxBatchLaunch L;
    L.rp = b.rp;
    L.tip = b.tip;
    L.items = b.d_ritems;
    L.prefix = b.d_tprefix;
    L.nitems = b.nitems;
    L.ntiles = b.ntiles;
    L.grids = b.grids;
    L.f64 = b.f64;
    L.queue = b.ctl;
    L.pos = b.pos;
    L.force = b.force;
    L.iters = b.iters;
    L.status = b.status;
    L.counters = b.counters();
    L.tile_y = b.tuning.tile_y;
    L.waves_per_block = b.tuning.waves_per_block;
    L.nblocks = 1;
    return L;
}

// Step-0 estimates of the first nitems items (all: b.nitems).
int enqueue_predict(ppafm_batch &b, int nitems = -1)
{
    if (!b.warm.enable || !b.ws) return 0;
    RelaxBatchLaunch L = base_launch(b);
    if (nitems >= 0) L.nitems = nitems;
    return launch_ws_predict(ws_view(b.ws), L, b.d_wprefix, b.wprefix[L.nitems], b.pred, b.warm.net_passes, b.st);
}

int enqueue_df(ppafm_batch &b)
{
    const int nw = (int)b.w.size();
    for (const DfGroup &g : b.groups) {
        int e = launch_convolve_z(g.ncol, g.nz, g.nz * 4, 4, 2, (const float *)(b.force + g.pos_off), nw, b.w.data(), b.df_scale,
                                  b.df + g.df_off, 0, b.st);
        if (e) return e;
    }
    return 0;
}

// stages: 1 grid, 2 relaxation (step-0 estimates first when warm starts
// with a model are on), 4 df, 8 estimates only, 16 FIRE only.
int enqueue_stages(ppafm_batch &b, int stages)
{
    int e = 0;
    if (stages & (2 | 16)) {
        e = launch_fill_u32(b.ctl, 16, 0u, b.st);
        if (e) return e;
    }
    if (stages & 1) {
        e = enqueue_grid(b, b.nitems, b.brick_prefix[b.nitems]);
        if (e) return e;
    }
    if (stages & (2 | 8)) {
        e = enqueue_predict(b);
        if (e) return e;
    }
    if (stages & (2 | 16)) {
        e = enqueue_relax(b, b.tuning, b.d_ritems, b.d_tprefix, b.nitems, b.ntiles, b.ritems);
        if (e) return e;
    }
    if ((stages & 4) && !b.w.empty()) {
        e = enqueue_df(b);
        if (e) return e;
    }
    return 0;
}

void drop_graph(ppafm_batch &b)
{
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
}

int capture(ppafm_batch &b)
{
    drop_graph(b);
    PP_TRY_HIP(hipStreamBeginCapture(b.st, hipStreamCaptureModeThreadLocal));
    const int e = enqueue_stages(b, 7);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b.st, &g);
    if (e) {
        if (g) (void)hipGraphDestroy(g);
        return e;
    }
    if (ee != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
        return -1;
    }
    b.graph = g;
    PP_TRY_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    return 0;
}

// Main relaxation table for the plan's tuning, then a fresh graph.
int apply_tuning(ppafm_batch &b)
{
    std::vector<BatchRelaxItem> it;
    std::vector<long long> prefix;
    relax_tables(b, b.nitems, b.tuning.tile_y, it, prefix);
    b.homog = homogeneous(it);
    PP_TRY_HIP(hipStreamSynchronize(b.st));  // no run may still read the old tables
    drop_graph(b);
    if (b.d_ritems) { (void)hipFree(b.d_ritems); b.d_ritems = nullptr; }
    if (b.d_tprefix) { (void)hipFree(b.d_tprefix); b.d_tprefix = nullptr; }
    int e = upload(b.d_ritems, it, b.st);
    if (e) return e;
    e = upload(b.d_tprefix, prefix, b.st);
    if (e) return e;
    b.ntiles = prefix[b.nitems];
    b.ritems = std::move(it);
    PP_TRY_HIP(hipStreamSynchronize(b.st));
    return capture(b);
}

int check_tuning(const ppafm_batch_tuning &t)
{
    PP_REQUIRE(t.tile_y == 1 || t.tile_y == 2 || t.tile_y == 4 || t.tile_y == 8 || t.tile_y == 16,
               "tile_y must be 1, 2, 4, 8 or 16");
    PP_REQUIRE(t.waves_per_block >= 1 && t.waves_per_block <= 4, "waves_per_block must be 1..4");
    PP_REQUIRE(t.blocks_per_cu >= 1 && t.blocks_per_cu <= 64, "blocks_per_cu must be 1..64");
    PP_REQUIRE(t.refill >= 0 && t.refill <= 3, "refill must be 0, 1, 2 or 3");
    return 0;
}

void free_plan(ppafm_batch *b)
{
    drop_graph(*b);
    void *bufs[] = {b->grids, b->pos, b->force, b->iters, b->df, b->atoms, b->ctl, b->d_gitems,
                    b->d_bprefix, b->d_ritems, b->d_tprefix, b->d_sitems, b->d_sprefix, b->pred, b->d_wprefix,
                    b->d_pprefix, b->status};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (b->own_stream && b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

int build_plan(ppafm_batch *b, int nitems, const ppafm_batch_item *items)
{
    PP_REQUIRE(nitems >= 1, "a batch needs at least one item");
    PP_REQUIRE(items, "items must not be NULL");
    b->nitems = nitems;
    b->items.assign(items, items + nitems);
    b->grid_off.assign(nitems + 1, 0);
    b->pos_off.assign(nitems + 1, 0);
    b->df_off.assign(nitems + 1, 0);
    b->atom_off.assign(nitems + 1, 0);
    b->brick_prefix.assign(nitems + 1, 0);
    b->gitems.resize(nitems);
    const int nw = (int)b->w.size();
    for (int k = 0; k < nitems; k++) {
        const ppafm_batch_item &s = items[k];
        PP_REQUIRE(s.n[0] > 0 && s.n[1] > 0 && s.n[2] > 0, "grid dimensions must be positive");
        PP_REQUIRE((long long)s.n[0] * s.n[1] * s.n[2] < (1LL << 31), "grid too large for 32-bit indices");
        PP_REQUIRE(s.natoms >= 0, "natoms must be >= 0");
        PP_REQUIRE(s.scan.nx >= 0 && s.scan.ny >= 0 && s.scan.nz >= 0, "scan sizes must be >= 0");
        const long long npts = (long long)s.n[0] * s.n[1] * s.n[2];
        const long long npos = (long long)s.scan.nx * s.scan.ny * s.scan.nz;
        const int nzout = nw > 0 ? s.scan.nz - nw + 1 : 0;
        b->grid_off[k + 1] = b->grid_off[k] + npts;
        b->pos_off[k + 1] = b->pos_off[k] + npos;
        b->df_off[k + 1] = b->df_off[k] + (nzout > 0 ? (long long)s.scan.nx * s.scan.ny * nzout : 0);
        b->atom_off[k + 1] = b->atom_off[k] + s.natoms;
        b->brick_prefix[k + 1] = b->brick_prefix[k] + ff_bricks(s.n);
        BatchGridItem &g = b->gitems[k];
        std::memset(&g, 0, sizeof(g));
        GridGeom gg = make_geom(s.n, s.cell);
        for (int a = 0; a < 3; a++) g.n[a] = s.n[a];
        g.natoms = s.natoms;
        g.atom_off = b->atom_off[k];
        g.grid_off = b->grid_off[k];
        for (int a = 0; a < 9; a++) g.dcell[a] = gg.dcell[a];
    }
    PP_REQUIRE(b->brick_prefix[nitems] < (1LL << 31), "batch has too many grid bricks for one launch");
    // df groups: runs of equal nz among items with nz >= nw
    if (nw > 0) {
        for (int k = 0; k < nitems;) {
            const int nz = items[k].scan.nz;
            int e = k + 1;
            while (e < nitems && items[e].scan.nz == nz) e++;
            if (nz - nw + 1 > 0) {
                DfGroup g;
                g.first = k; g.last = e; g.nz = nz;
                g.ncol = (b->pos_off[e] - b->pos_off[k]) / nz;
                g.pos_off = b->pos_off[k];
                g.df_off = b->df_off[k];
                if (g.ncol > 0) b->groups.push_back(g);
            }
            k = e;
        }
    }
    const size_t esz = b->f64 ? 32 : 16;
    const long long npos = b->pos_off[nitems];
    PP_TRY_HIP(hipMalloc(&b->grids, std::max<long long>(b->grid_off[nitems], 1) * esz));
    PP_TRY_HIP(hipMalloc((void **)&b->pos, std::max<long long>(npos, 1) * 12));
    PP_TRY_HIP(hipMalloc((void **)&b->force, std::max<long long>(npos, 1) * 16));
    if (b->keep_iters) {
        PP_TRY_HIP(hipMalloc((void **)&b->iters, std::max<long long>(npos, 1) * 4));
        PP_TRY_HIP(hipMalloc((void **)&b->status, std::max<long long>(npos, 1)));
    }
    PP_TRY_HIP(hipMalloc((void **)&b->df, std::max<long long>(b->df_off[nitems], 1) * 4));
    b->atom_cap = std::max<long long>(b->atom_off[nitems], 1);
    PP_TRY_HIP(hipMalloc((void **)&b->atoms, b->atom_cap * sizeof(AtomRec)));
    PP_TRY_HIP(hipMalloc((void **)&b->ctl, 64));
    PP_TRY_HIP(hipMemsetAsync(b->atoms, 0, std::max<long long>(b->atom_off[nitems], 1) * sizeof(AtomRec), b->st));
    PP_TRY_HIP(hipMemsetAsync(b->ctl, 0, 64, b->st));
    int e = upload(b->d_gitems, b->gitems, b->st);
    if (e) return e;
    e = upload(b->d_bprefix, b->brick_prefix, b->st);
    if (e) return e;
    e = upload(b->d_pprefix, b->pos_off, b->st);
    if (e) return e;
    return apply_tuning(*b);
}

}  // namespace

extern "C" {

ppafm_batch *ppafm_batch_create(int nitems, const ppafm_batch_item *items, const ppafm_ff_model *model,
                                const ppafm_relax_params *rp, const ppafm_tip *tip, int grid_f64, int nw,
                                const double *df_w, double df_scale, int keep_iters,
                                const ppafm_batch_tuning *tuning, void *stream)
{
    if (!model || !rp || !tip) {
        set_error("model, rp and tip must not be NULL");
        return nullptr;
    }
    if (nw < 0 || nw > 64 || (nw > 0 && !df_w)) {
        set_error("df weights: need 0 <= nw <= 64 and a weight array");
        return nullptr;
    }
    if (rp->alg != 1 || rp->max_iters < 1) {
        set_error("batched relaxation needs FIRE (alg 1) and max_iters >= 1");
        return nullptr;
    }
    ppafm_batch *b = new ppafm_batch;
    b->model = *model;
    b->rp = *rp;
    b->tip = *tip;
    b->f64 = grid_f64 != 0;
    b->w.assign(df_w, df_w + nw);
    b->df_scale = df_scale;
    b->keep_iters = keep_iters != 0;
    if (tuning) {
        if (check_tuning(*tuning)) {
            delete b;
            return nullptr;
        }
        b->tuning = *tuning;
    }
    if (stream) {
        b->st = as_stream(stream);
    } else {
        if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
            set_error("hipStreamCreateWithFlags failed");
            delete b;
            return nullptr;
        }
        b->own_stream = true;
    }
    if (device_cus(b->ncu) || build_plan(b, nitems, items)) {
        free_plan(b);
        return nullptr;
    }
    return b;
}

int ppafm_batch_set_atoms(ppafm_batch *b, const ppafm_atom *atoms)
{
    PP_REQUIRE(b, "batch is NULL");
    const long long n = b->atom_off[b->nitems];
    PP_REQUIRE(atoms || n == 0, "atoms must not be NULL");
    if (n > 0) PP_TRY_HIP(hipMemcpyAsync(b->atoms, atoms, n * sizeof(AtomRec), hipMemcpyHostToDevice, b->st));
    PP_TRY_HIP(hipStreamSynchronize(b->st));
    // the grid kernel skips uncharged atoms' Coulomb terms (exact zeros) when
    // there are any; a change of that re-captures the graph
    bool any = false;
    for (long long k = 0; k < n && !any; k++) any = atoms[k].q == 0.0;
    if (any != b->any_uncharged || !b->exec) {
        b->any_uncharged = any;
        return capture(*b);
    }
    return 0;
}

int ppafm_batch_set_atoms_n(ppafm_batch *b, const int *natoms, const ppafm_atom *atoms)
{
    PP_REQUIRE(b && natoms, "batch or natoms is NULL");
    long long total = 0;
    for (int k = 0; k < b->nitems; k++) {
        PP_REQUIRE(natoms[k] >= 0, "natoms must be >= 0");
        total += natoms[k];
    }
    PP_REQUIRE(atoms || total == 0, "atoms must not be NULL");
    PP_TRY_HIP(hipStreamSynchronize(b->st));
    bool recapture = false;
    if (total > b->atom_cap) {
        // a larger buffer: the graph holds the old address, so it goes
        // before the old buffer is freed; an error from here on leaves the
        // plan without a graph (ppafm_batch_run refuses it), never with one
        // that reads freed memory, and the next successful call captures
        const long long cap = total + total / 4;
        AtomRec *p = nullptr;
        PP_TRY_HIP(hipMalloc((void **)&p, cap * sizeof(AtomRec)));
        drop_graph(*b);
        if (b->atoms) (void)hipFree(b->atoms);
        b->atoms = p;
        b->atom_cap = cap;
        recapture = true;
    }
    for (int k = 0; k < b->nitems; k++) {
        b->items[k].natoms = natoms[k];
        b->atom_off[k + 1] = b->atom_off[k] + natoms[k];
        b->gitems[k].natoms = natoms[k];
        b->gitems[k].atom_off = b->atom_off[k];
    }
    // the grid kernel reads each item's atom count and offset from this
    // table when it runs
    int e = upload(b->d_gitems, b->gitems, b->st);
    if (e) return e;
    if (total > 0) PP_TRY_HIP(hipMemcpyAsync(b->atoms, atoms, total * sizeof(AtomRec), hipMemcpyHostToDevice, b->st));
    PP_TRY_HIP(hipStreamSynchronize(b->st));
    bool any = false;
    for (long long k = 0; k < total && !any; k++) any = atoms[k].q == 0.0;
    if (any != b->any_uncharged) {
        b->any_uncharged = any;
        recapture = true;
    }
    // a plan always has a graph once created: capture when this call needs a
    // new one or an earlier failed call dropped it
    if (recapture || !b->exec) return capture(*b);
    return 0;
}

int ppafm_batch_run(ppafm_batch *b)
{
    PP_REQUIRE(b && b->exec, "batch has no graph");
    PP_TRY_HIP(hipGraphLaunch(b->exec, b->st));
    return 0;
}

int ppafm_batch_run_stages(ppafm_batch *b, int stages)
{
    PP_REQUIRE(b, "batch is NULL");
    PP_REQUIRE(stages >= 0 && stages <= 31, "stages is a mask of bits 0..4");
    return enqueue_stages(*b, stages);
}

int ppafm_batch_tune(ppafm_batch *b, int nsample, int reps, ppafm_batch_tuning *best, double *best_ms)
{
    PP_REQUIRE(b, "batch is NULL");
    PP_REQUIRE(reps >= 1, "reps must be >= 1");
    const int ns = std::max(1, std::min(nsample, b->nitems));
    // candidates are timed on the relaxation the plan runs (warm starts
    // included); grids and step-0 estimates of the sample, once
    int e = launch_fill_u32(b->ctl, 12, 0u, b->st);
    if (e) return e;
    e = enqueue_grid(*b, ns, b->brick_prefix[ns]);
    if (e) return e;
    e = enqueue_predict(*b, ns);
    if (e) return e;
    static const int tys[] = {2, 4, 8, 16};
    static const int wpbs[] = {1, 2, 4};
    static const int wpcs[] = {8, 16, 32};
    hipEvent_t e0, e1;
    PP_TRY_HIP(hipEventCreate(&e0));
    PP_TRY_HIP(hipEventCreate(&e1));
    double best_t = 1e300;
    ppafm_batch_tuning win = b->tuning;
    int rc = 0;
    for (int ty : tys) {
        std::vector<BatchRelaxItem> it;
        std::vector<long long> prefix;
        relax_tables(*b, ns, ty, it, prefix);
        if (b->d_sitems) { (void)hipFree(b->d_sitems); b->d_sitems = nullptr; }
        if (b->d_sprefix) { (void)hipFree(b->d_sprefix); b->d_sprefix = nullptr; }
        if ((rc = upload(b->d_sitems, it, b->st)) || (rc = upload(b->d_sprefix, prefix, b->st))) break;
        // lane refill only for one shared geometry: the tile queue and one
        // column at a time (wave pools, refill 2 and 3, measured slower than
        // the tile queue on every shape, are not candidates)
        const int nrf = homogeneous(it) ? 2 : 1;
        for (int rf = 0; rf < nrf; rf++)
        for (int wpb : wpbs)
            for (int wpc : wpcs) {
                const ppafm_batch_tuning t{ty, wpb, std::max(1, wpc / wpb), rf};
                double tot = 0;
                for (int r = 0; r <= reps && rc == 0; r++) {
                    if ((rc = launch_fill_u32(b->ctl, 12, 0u, b->st))) break;
                    if (hipEventRecord(e0, b->st) != hipSuccess) { rc = -1; break; }
                    if ((rc = enqueue_relax(*b, t, b->d_sitems, b->d_sprefix, ns, prefix[ns], it))) break;
                    if (hipEventRecord(e1, b->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                        set_error("event failure while tuning");
                        rc = -1;
                        break;
                    }
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (r > 0) tot += ms;  // first run warms up
                }
                if (rc) break;
                const double avg = tot / reps;
                if (avg < best_t) {
                    best_t = avg;
                    win = t;
                }
            }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        if (rc == -2 || rc == -1) return rc;
        return -1;
    }
    b->tuning = win;
    if (best) *best = win;
    if (best_ms) *best_ms = best_t;
    return apply_tuning(*b);
}

int ppafm_batch_set_tuning(ppafm_batch *b, const ppafm_batch_tuning *tuning)
{
    PP_REQUIRE(b && tuning, "batch and tuning must not be NULL");
    int e = check_tuning(*tuning);
    if (e) return e;
    b->tuning = *tuning;
    return apply_tuning(*b);
}

int ppafm_batch_get_tuning(const ppafm_batch *b, ppafm_batch_tuning *tuning)
{
    PP_REQUIRE(b && tuning, "batch and tuning must not be NULL");
    *tuning = b->tuning;
    return 0;
}

void *ppafm_batch_buffer(const ppafm_batch *b, int which)
{
    if (!b) return nullptr;
    switch (which) {
    case 0: return b->grids;
    case 1: return b->pos;
    case 2: return b->force;
    case 3: return b->iters;
    case 4: return b->df;
    case 5: return b->counters();
    case 6: return b->atoms;
    case 7: return b->pred;
    case 8: return b->status;
    default: return nullptr;
    }
}

int ppafm_batch_offsets(const ppafm_batch *b, long long *grid_off, long long *pos_off, long long *df_off)
{
    PP_REQUIRE(b, "batch is NULL");
    const size_t n = (size_t)b->nitems + 1;
    if (grid_off) std::copy(b->grid_off.begin(), b->grid_off.begin() + n, grid_off);
    if (pos_off) std::copy(b->pos_off.begin(), b->pos_off.begin() + n, pos_off);
    if (df_off) std::copy(b->df_off.begin(), b->df_off.begin() + n, df_off);
    return 0;
}

void *ppafm_batch_stream(const ppafm_batch *b) { return b ? (void *)b->st : nullptr; }

int ppafm_batch_set_warmstart(ppafm_batch *b, const ppafm_warm_params *p, ppafm_ws_model *m)
{
    PP_REQUIRE(b && p, "batch and parameters must not be NULL");
    if (p->enable) {
        PP_REQUIRE(p->newton_max >= 0 && p->newton_max <= 64, "newton_max must be 0..64");
        PP_REQUIRE(p->newton_max0 >= 0 && p->newton_max0 <= 64, "newton_max0 must be 0..64");
        PP_REQUIRE(p->newton_retry >= 0 && p->newton_retry <= 1000, "newton_retry must be 0..1000");
        PP_REQUIRE(p->newton_tol > 0 && p->newton_tol < 1, "newton_tol must be in (0, 1) A");
        PP_REQUIRE(p->newton_cap > 0 && p->newton_cap <= 1, "newton_cap must be in (0, 1] A");
        PP_REQUIRE(p->extrap_max >= 0 && p->extrap_max <= 1, "extrap_max must be in [0, 1] A");
        PP_REQUIRE(p->residual_max >= 0 && p->residual_max <= 1e6, "residual_max must be in [0, 1e6] eV/A");
        PP_REQUIRE(!m || (p->net_passes >= 1 && p->net_passes <= 8), "net_passes must be 1..8");
    }
    PP_TRY_HIP(hipStreamSynchronize(b->st));
    if (p->enable && m && !b->pred) {
        std::vector<long long> wp(b->nitems + 1, 0);
        long long ncol = 0;
        for (int k = 0; k < b->nitems; k++) {
            const ppafm_batch_item &it = b->items[k];
            const long long nc = it.scan.nz > 0 ? (long long)it.scan.nx * it.scan.ny : 0;
            const long long tiles = nc > 0 ? (long long)((it.scan.nx + 7) / 8) * ((it.scan.ny + 7) / 8) : 0;
            wp[k + 1] = wp[k] + tiles;  // 8x8 pixel tiles of the step-0 estimates
            ncol += nc;
        }
        PP_REQUIRE(ncol < (1LL << 31), "too many pixel columns for warm starts");
        PP_TRY_HIP(hipMalloc((void **)&b->pred, std::max<long long>(ncol, 1) * 12));
        b->nwaves = wp[b->nitems];
        b->wprefix = wp;
        int e = upload(b->d_wprefix, wp, b->st);
        if (e) return e;
        PP_TRY_HIP(hipStreamSynchronize(b->st));
    }
    b->warm = *p;
    b->ws = p->enable ? m : nullptr;
    return capture(*b);
}

int ppafm_batch_polish(ppafm_batch *b, double conv_f2, int max_iters, float *out_dev, int *iters_dev)
{
    PP_REQUIRE(b && out_dev, "batch and output must not be NULL");
    PP_REQUIRE(max_iters >= 1, "max_iters must be >= 1");
    RelaxBatchLaunch L = base_launch(*b);
    L.rp.conv_f2 = conv_f2;
    L.rp.max_iters = max_iters;
    return launch_fire_points(L, b->d_pprefix, b->pos_off[b->nitems], b->pos, out_dev, iters_dev, b->st);
}

int ppafm_batch_ws_samples(ppafm_batch *b, const float *truth_dev, long long nsel, const long long *sel_dev,
                           unsigned short *patches_dev, float *targets_dev, double jitter, unsigned seed)
{
    PP_REQUIRE(b && truth_dev && (nsel == 0 || (sel_dev && patches_dev && targets_dev)), "NULL argument");
    RelaxBatchLaunch L = base_launch(*b);
    return launch_ws_samples(L, b->d_pprefix, b->pos, truth_dev, nsel, sel_dev, patches_dev, targets_dev, (float)jitter,
                             seed, b->st);
}

void ppafm_batch_destroy(ppafm_batch *b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->st);
    free_plan(b);
}

}  // extern "C"
