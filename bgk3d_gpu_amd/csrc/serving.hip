// Continuous-batching server (bgks_* in include/bgk3d_engine.h): many
// independent lid-driven cavity solves of different sizes, viscosities and
// storage types advanced together on one GPU.
//
// The reference runs such a sweep as separate programs, one after the other
// (UTIL/run_sweep.sh:55-63, 128-169), each with its own time loop
// (SRC/bgk3d.F90:169-196).  Here every admitted request is an engine handle
// whose two lattices live in one device arena sized by the caller's HBM
// budget; requests are admitted in arrival order while the arena can hold
// them.  The requests of a chunk are split by (storage type, rounding, row
// layout) group into sub-batches of at most 512 MB of lattices, and each
// sub-batch runs all the chunk's steps, one launch per step, before the
// next.  A request alone in its sub-batch runs the engine's own kernel; the
// launch grid of a larger sub-batch is the concatenation of its requests'
// row blocks, and a block finds its request through a per-block owner table
// and then runs the same per-cell code as the single-request kernel.  So a
// request's result does not depend on what it shares the GPU with
// (stochastic-rounding streams are keyed by cell index, request seed and
// the request's own step count).
//
// Steps run in chunks of at most `chunk_steps`, shortened so that no request
// overshoots its step count and, with a chunk-time target, so that a chunk
// takes about that long (requests join and leave only between chunks, so
// the target bounds how long an arrival waits to start).  Every cell's new
// rest population is checked for a non-finite
// value inside the step kernel; a request whose solution became non-finite
// is retired as unstable at the end of the chunk instead of running to its
// last step.
// Policies: every admitted request advances in every chunk (processor
// sharing, arrival-order admission), or only the admitted requests with the
// least work left, as many as keep a given number of cells per step in
// flight (shortest remaining work first, which cuts waiting for the small
// requests of a mix while the GPU stays saturated); a request passed over
// for kMaxSkips chunks in a row is advanced (and a queued one admitted)
// ahead of the rest, so a stream of small requests cannot starve a large
// one.
#include "engine_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

using namespace bgk;
using namespace bgk::eng;

namespace {

// Block = 384 threads for every request: rows of 1, 2, 3 or 6 waves (up to
// 384 cells) fill a block with no idle wave (6, 3, 2 or 1 rows per block).
constexpr int kServeWaves = 6;
constexpr size_t kArenaAlign = 64 << 10;
// A chunk runs its steps sub-batch by sub-batch: consecutive requests of a
// storage group whose lattices add up to at most this many bytes (one larger
// request forms its own), each sub-batch all K steps before the next.  The
// fused kernel streams faster over a small working set, whose address
// translations stay cached from one launch to the next (measured on the
// 16-request cavity mix, FP32 part: 30.9k MLUPS with 512 MB sub-batches,
// 28.1k with 3 GB).  Every request advances the chunk's steps in order, so
// results do not depend on the grouping.  A request alone in its sub-batch
// runs the engine's own kernel (no block-to-request lookup); batching in one
// launch pays only for requests too small to fill the GPU by themselves.
constexpr size_t kSubBatchBytes = size_t(1) << 29;
// Requests admitted at once: each holds one non-finite flag word.
constexpr int kMaxFlags = 4096;
// Shortest-first aging: chunks a request may be passed over in a row.
constexpr int kMaxSkips = 8;

template <typename S>
struct BatchSlot {
    FusedArgs<S> args[2];  // indexed by the lattice the step reads
    RowMap rm;
    int gy;                // row blocks along j (padded rows)
    int nbx;               // position blocks per plane (compact rows)
    int par0;              // lattice read at chunk step 0
    uint32_t cfg_seed;
    long long steps0;      // steps completed before the chunk
    uint32_t block0;       // first block of this request in the launch
};

// One time step of every request of one storage group.  The lid-driven
// cavity: bounce-back walls in x and y, lid at z = n + 1 (SRC/bcond.driven.F90).
template <typename S, int SR>
__global__ __launch_bounds__(64 * kServeWaves) __attribute__((amdgpu_waves_per_eu(FusedWaves<S, 0>::value))) void
collide_batched(const BatchSlot<S>* __restrict__ slots, const uint32_t* __restrict__ owner, uint32_t base, int t)
{
    const uint32_t b = base + blockIdx.x;
    const BatchSlot<S>& sl = slots[owner[b]];
    const RowMap rm = sl.rm;
    const uint32_t lb = b - sl.block0;
    const int bx = (int)(lb % (uint32_t)rm.bpr);
    const uint32_t rest = lb / (uint32_t)rm.bpr;
    const int by = (int)(rest % (uint32_t)sl.gy);
    const int bz = (int)(rest / (uint32_t)sl.gy);
    const uint32_t seed = SR ? step_seed(sl.cfg_seed, sl.steps0 + t) : 0u;
    fused_rows<S, AX_WALL, AX_WALL, AX_LID, false, 0, SR>(sl.args[(sl.par0 + t) & 1], rm, bx, by, bz, seed);
}

// The same for requests in the compact layout (collide_compact's mapping,
// 256-thread blocks): a request's blocks are nbx position blocks per plane.
template <typename S, int SR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FusedWaves<S, 0>::value))) void
collide_batched_compact(const BatchSlot<S>* __restrict__ slots, const uint32_t* __restrict__ owner, uint32_t base,
                        int t)
{
    const uint32_t b = base + blockIdx.x;
    const BatchSlot<S>& sl = slots[owner[b]];
    const uint32_t lb = b - sl.block0;
    const int bx = (int)(lb % (uint32_t)sl.nbx);
    const int bz = (int)(lb / (uint32_t)sl.nbx);
    const uint32_t seed = SR ? step_seed(sl.cfg_seed, sl.steps0 + t) : 0u;
    compact_cells<S, AX_WALL, AX_WALL, AX_LID, false, 0, SR>(sl.args[(sl.par0 + t) & 1], bx, bz, seed);
}

double clock_now()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) with coalescing of free ranges.
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length

    void reset(size_t n)
    {
        size = n;
        free_.clear();
        free_[0] = n;
    }
    bool alloc(size_t len, size_t& off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            off = it->first;
            const size_t rest = it->second - len;
            free_.erase(it);
            if (rest) free_[off + len] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t len)
    {
        auto it = free_.emplace(off, len).first;
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
    size_t used() const
    {
        size_t f = 0;
        for (auto& kv : free_) f += kv.second;
        return size - f;
    }
};

struct Instance {
    long long id = 0;
    bgks_request req{};
    bgk_handle h;
    int state = BGKS_QUEUED;
    bool has_mem = false;
    size_t off = 0, len = 0;
    int flag = -1;      // non-finite flag word while admitted
    int skipped = 0;    // chunks passed over in a row (queued or active)
    double diag[5] = {0, 0, 0, 0, 0};
    int unstable = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

// Launch group of a request: storage type and rounding (which select the
// kernel's template), plus 8 for the compact layout.
int group_of(int storage, int rounding, bool compact)
{
    const bool sr = rounding && (storage == BGK_F16 || storage == BGK_BF16);
    return storage * 2 + (sr ? 1 : 0) + (compact ? 8 : 0);
}
constexpr int kGroups = 16;

bgk_config request_config(const bgks_request& r, int device)
{
    bgk_config c;
    bgk_config_default(&c, r.n, r.n, r.n);
    c.boundary = BGK_BC_DRIVEN;
    c.storage = r.storage;
    c.rounding = r.rounding;
    c.svisc = r.svisc;
    c.u0 = r.u0;
    c.device = device;
    c.seed = r.seed;
    return c;
}

}  // namespace

struct bgks_server {
    int device = 0;
    int chunk = 100;
    double chunk_ms = 0;          // chunk-time target (0: chunk_steps only)
    double ns_per_cell_step = 0;  // measured by the last chunk
    hipStream_t stream = nullptr;
    uint32_t* d_flags = nullptr;  // non-finite flags, one word per admitted request
    uint32_t* h_flags = nullptr;
    std::vector<int> free_flags;
    void* arena_mem = nullptr;
    Arena arena;
    double* scratch = nullptr;
    double* hscratch = nullptr;
    std::deque<long long> queue;
    std::map<long long, std::unique_ptr<Instance>> all;
    long long next_id = 1;
    int policy = BGKS_ALL_ACTIVE;
    long long min_cells = 0;  // BGKS_SHORTEST_FIRST: cells advanced per step at least
    // per storage group: device tables and pinned staging
    struct Group {
        void* d_slots = nullptr;
        uint32_t* d_owner = nullptr;
        void* h_slots = nullptr;
        uint32_t* h_owner = nullptr;
        size_t slot_bytes_cap = 0, owner_cap = 0;
    } groups[kGroups];
    long long launches = 0, cell_steps = 0, chunks = 0;
};

namespace {

int grow_group(bgks_server& s, bgks_server::Group& g, size_t slot_bytes, size_t owners)
{
    if (slot_bytes > g.slot_bytes_cap) {
        if (g.d_slots) (void)hipFree(g.d_slots);
        if (g.h_slots) (void)hipHostFree(g.h_slots);
        g.d_slots = g.h_slots = nullptr;
        const size_t cap = std::max(slot_bytes, 2 * g.slot_bytes_cap);
        ENG_HIP(hipMalloc(&g.d_slots, cap));
        ENG_HIP(hipHostMalloc(&g.h_slots, cap, hipHostMallocDefault));
        g.slot_bytes_cap = cap;
    }
    if (owners > g.owner_cap) {
        if (g.d_owner) (void)hipFree(g.d_owner);
        if (g.h_owner) (void)hipHostFree(g.h_owner);
        g.d_owner = g.h_owner = nullptr;
        const size_t cap = std::max(owners, 2 * g.owner_cap);
        ENG_HIP(hipMalloc(&g.d_owner, cap * sizeof(uint32_t)));
        ENG_HIP(hipHostMalloc(&g.h_owner, cap * sizeof(uint32_t), hipHostMallocDefault));
        g.owner_cap = cap;
    }
    (void)s;
    return 0;
}

// Build and upload the launch tables of one storage group: a slot per
// request (kernel arguments for both lattice parities, row mapping, seed,
// step count at the chunk start) and the owner of every block.
struct SubBatch {
    uint32_t block0, nblocks;
    int single;  // the member that is alone in this sub-batch, else -1
};

template <typename S>
int prepare_group(bgks_server& s, bgks_server::Group& g, const std::vector<Instance*>& members,
                  std::vector<SubBatch>& subs, hipStream_t stream)
{
    std::vector<BatchSlot<S>> slots(members.size());
    uint32_t total = 0;
    size_t sub_bytes = 0;
    subs.clear();
    for (size_t q = 0; q < members.size(); ++q) {
        bgk_handle& h = members[q]->h;
        if (subs.empty() || (sub_bytes > 0 && sub_bytes + members[q]->len > kSubBatchBytes)) {
            subs.push_back({total, 0, (int)q});
            sub_bytes = 0;
        } else {
            subs.back().single = -1;
        }
        sub_bytes += members[q]->len;
        BatchSlot<S>& sl = slots[q];
        sl.args[0] = fused_args<S>(h, 0);
        sl.args[1] = fused_args<S>(h, 1);
        sl.args[0].bad = sl.args[1].bad = s.d_flags + members[q]->flag;
        sl.rm = row_map_capped(h.cfg.l, kServeWaves);
        sl.gy = ceil_div(h.cfg.m, sl.rm.rpb);
        sl.nbx = ceil_div(sl.args[0].lend, 256);
        sl.par0 = h.cur;
        sl.cfg_seed = h.cfg.seed;
        sl.steps0 = h.steps;
        sl.block0 = total;
        const long long nb = h.compact ? (long long)sl.nbx * h.cfg.n : (long long)sl.rm.bpr * sl.gy * h.cfg.n;
        if ((long long)total + nb > 0xffffffffLL) return fail(BGK_EINVAL, "batch grid exceeds 2^32 blocks");
        if ((long long)subs.back().nblocks + nb > 0x7fffffffLL)
            return fail(BGK_EINVAL, "sub-batch grid exceeds 2^31 blocks");
        total += (uint32_t)nb;
        subs.back().nblocks += (uint32_t)nb;
    }
    if (int rc = grow_group(s, g, slots.size() * sizeof(BatchSlot<S>), total)) return rc;
    std::memcpy(g.h_slots, slots.data(), slots.size() * sizeof(BatchSlot<S>));
    for (size_t q = 0; q < members.size(); ++q) {
        const uint32_t b0 = slots[q].block0;
        const uint32_t b1 = q + 1 < members.size() ? slots[q + 1].block0 : total;
        std::fill(g.h_owner + b0, g.h_owner + b1, (uint32_t)q);
    }
    ENG_HIP(hipMemcpyAsync(g.d_slots, g.h_slots, slots.size() * sizeof(BatchSlot<S>), hipMemcpyHostToDevice,
                           stream));
    ENG_HIP(hipMemcpyAsync(g.d_owner, g.h_owner, (size_t)total * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    return 0;
}

template <typename S, int SR>
void launch_group(bgks_server::Group& g, const SubBatch& sb, int t, hipStream_t stream, bool compact)
{
    if (sb.single >= 0) {
        // the engine's launch for a driven cavity (engine.hip launch_fused)
        const BatchSlot<S>& sl = static_cast<const BatchSlot<S>*>(g.h_slots)[sb.single];
        FusedArgs<S> p = sl.args[(sl.par0 + t) & 1];
        p.seed = step_seed(sl.cfg_seed, sl.steps0 + t);
        if (compact) {
            hipLaunchKernelGGL((collide_compact<S, AX_WALL, AX_WALL, AX_LID, false, 0, SR>),
                               dim3((unsigned)ceil_div(p.lend, 256), (unsigned)p.n), dim3(256), compact_lds_cap<S>(),
                               stream, p);
            return;
        }
        const RowMap rm = row_map(p.l);
        if constexpr (sizeof(S) == 2) {
            if (p.lend <= p.l) {  // no line-completing lanes: the form without the selects
                hipLaunchKernelGGL((collide_fused<S, AX_WALL, AX_WALL, AX_LID, false, 0, SR, false>),
                                   fused_grid(rm, p.m, p.n), fused_block(rm), 0, stream, p, rm);
                return;
            }
        }
        hipLaunchKernelGGL((collide_fused<S, AX_WALL, AX_WALL, AX_LID, false, 0, SR>), fused_grid(rm, p.m, p.n),
                           fused_block(rm), 0, stream, p, rm);
        return;
    }
    if (compact)
        hipLaunchKernelGGL((collide_batched_compact<S, SR>), dim3(sb.nblocks), dim3(256), compact_lds_cap<S>(), stream,
                           static_cast<const BatchSlot<S>*>(g.d_slots), g.d_owner, sb.block0, t);
    else
        hipLaunchKernelGGL((collide_batched<S, SR>), dim3(sb.nblocks), dim3(64 * kServeWaves), 0, stream,
                           static_cast<const BatchSlot<S>*>(g.d_slots), g.d_owner, sb.block0, t);
}

template <typename F>
int by_group(int gid, F&& f)
{
    switch (gid & 7) {
        case 0: return f(float{}, std::integral_constant<int, 0>{});
        case 2: return f(double{}, std::integral_constant<int, 0>{});
        case 4: return f(f16_t{}, std::integral_constant<int, 0>{});
        case 5: return f(f16_t{}, std::integral_constant<int, 1>{});
        case 6: return f(bf16_t{}, std::integral_constant<int, 0>{});
        case 7: return f(bf16_t{}, std::integral_constant<int, 1>{});
    }
    return fail(BGK_EINVAL, "storage group %d", gid);
}

long long remaining_work(const Instance& in)
{
    return (long long)in.h.cfg.l * in.h.cfg.m * in.h.cfg.n * (in.req.steps - in.h.steps);
}

int start(bgks_server& s, Instance& in, size_t off)
{
    in.off = off;
    in.has_mem = true;
    void* region = static_cast<char*>(s.arena_mem) + off;
    place_lattices(in.h, region);
    ENG_HIP(hipMemsetAsync(region, 0, in.len, s.stream));  // rest state: shifted populations are zero
    in.flag = s.free_flags.back();
    s.free_flags.pop_back();
    ENG_HIP(hipMemsetAsync(s.d_flags + in.flag, 0, sizeof(uint32_t), s.stream));
    in.state = BGKS_ACTIVE;
    in.skipped = 0;
    in.t_admit = clock_now();
    return 0;
}

// Arrival order: the head of the queue waits for room.  Shortest first:
// queued requests in order of their work, each admitted if it fits, except
// that requests passed over kMaxSkips times come first, in arrival order,
// and the first of them that does not fit stops admission (it waits for
// room instead of being overtaken again).
int admit(bgks_server& s)
{
    if (s.policy == BGKS_ALL_ACTIVE) {
        while (!s.queue.empty() && !s.free_flags.empty()) {
            Instance& in = *s.all.at(s.queue.front());
            size_t off = 0;
            if (!s.arena.alloc(in.len, off)) break;
            if (int rc = start(s, in, off)) return rc;
            s.queue.pop_front();
        }
        return 0;
    }
    std::vector<long long> order(s.queue.begin(), s.queue.end());
    std::stable_sort(order.begin(), order.end(), [&](long long a, long long b) {
        const Instance &x = *s.all.at(a), &y = *s.all.at(b);
        const bool ax = x.skipped >= kMaxSkips, ay = y.skipped >= kMaxSkips;
        if (ax != ay) return ax;
        if (ax) return false;  // both aged: arrival order
        return remaining_work(x) < remaining_work(y);
    });
    for (long long id : order) {
        if (s.free_flags.empty()) break;
        Instance& in = *s.all.at(id);
        size_t off = 0;
        if (!s.arena.alloc(in.len, off)) {
            if (in.skipped >= kMaxSkips) break;
            continue;
        }
        if (int rc = start(s, in, off)) return rc;
        s.queue.erase(std::find(s.queue.begin(), s.queue.end(), id));
    }
    return 0;
}

void free_lattice(bgks_server& s, Instance& in)
{
    if (in.flag >= 0) {
        s.free_flags.push_back(in.flag);
        in.flag = -1;
    }
    if (!in.has_mem) return;
    s.arena.release(in.off, in.len);
    in.has_mem = false;
    in.h.mem[0] = in.h.mem[1] = nullptr;
}

}  // namespace

namespace bgk {
namespace eng {

// Shortest remaining work first with aging: every request passed over
// max_skips chunks in a row, then the others in order of remaining work
// (stable) until min_cells cells per step are in flight (at least one).
std::vector<int> plan_shortest(const long long* remaining, const long long* cells, const int* skipped, int count,
                               long long min_cells, int max_skips)
{
    std::vector<int> order(count), out;
    for (int q = 0; q < count; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return remaining[a] < remaining[b]; });
    long long flight = 0;
    std::vector<char> taken(count, 0);
    for (int q = 0; q < count; ++q)
        if (skipped[q] >= max_skips) {
            out.push_back(q);
            taken[q] = 1;
            flight += cells[q];
        }
    for (int q : order) {
        if (taken[q]) continue;
        if (!out.empty() && flight >= min_cells) break;
        out.push_back(q);
        flight += cells[q];
    }
    return out;
}

}  // namespace eng
}  // namespace bgk

extern "C" {

double bgks_clock(void) { return clock_now(); }

int bgks_request_bytes(const bgks_request* r, size_t* bytes)
{
    if (!r || !bytes) return fail(BGK_EINVAL, "null argument");
    bgk_handle h;
    if (int rc = handle_setup(h, request_config(*r, 0))) return rc;
    *bytes = round_up((long)lattice_bytes(h), (long)kArenaAlign);
    return 0;
}

int bgks_create(int device, size_t budget_bytes, int chunk_steps, bgks_server** out)
{
    if (!out) return fail(BGK_EINVAL, "null argument");
    *out = nullptr;
    if (chunk_steps < 1) return fail(BGK_EINVAL, "chunk_steps %d", chunk_steps);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(BGK_EINVAL, "device %d not available (%d devices)", device, ndev);
    ENG_HIP(hipSetDevice(device));
    budget_bytes = budget_bytes / kArenaAlign * kArenaAlign;
    if (budget_bytes == 0) return fail(BGK_EINVAL, "budget below %zu bytes", kArenaAlign);
    auto* s = new bgks_server();
    s->device = device;
    s->chunk = chunk_steps;
    auto bail = [&](int code) {
        bgks_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(BGK_EHIP, "stream creation failed"));
    if (hipMalloc(&s->d_flags, kMaxFlags * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc(&s->h_flags, kMaxFlags * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return bail(fail(BGK_ENOMEM, "flags"));
    for (int q = kMaxFlags - 1; q >= 0; --q) s->free_flags.push_back(q);
    if (hipMalloc(&s->arena_mem, budget_bytes) != hipSuccess)
        return bail(fail(BGK_ENOMEM, "cannot reserve %.2f GB for the request arena", budget_bytes / 1e9));
    s->arena.reset(budget_bytes);
    if (hipMalloc(&s->scratch, (kRedBlocks + 8) * 5 * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&s->hscratch, (kRedBlocks + 8) * 5 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bail(fail(BGK_ENOMEM, "scratch"));
    *out = s;
    return 0;
}

int bgks_set_chunk_time(bgks_server* s, double target_ms)
{
    if (!s) return fail(BGK_EINVAL, "null server");
    if (!(target_ms >= 0.0)) return fail(BGK_EINVAL, "chunk time %g ms", target_ms);
    s->chunk_ms = target_ms;
    return 0;
}

int bgks_schedule(const long long* remaining, const long long* cells, const int* skipped, int count,
                  long long min_cells_per_step, int max_skips, int* selected, int* n_selected)
{
    if (count < 0 || (count > 0 && (!remaining || !cells || !skipped || !selected)) || !n_selected)
        return fail(BGK_EINVAL, "bad arguments");
    const std::vector<int> sel = plan_shortest(remaining, cells, skipped, count, min_cells_per_step, max_skips);
    for (size_t q = 0; q < sel.size(); ++q) selected[q] = sel[q];
    *n_selected = (int)sel.size();
    return 0;
}

int bgks_set_policy(bgks_server* s, int policy, long long min_cells_per_step)
{
    if (!s) return fail(BGK_EINVAL, "null server");
    if (policy != BGKS_ALL_ACTIVE && policy != BGKS_SHORTEST_FIRST) return fail(BGK_EINVAL, "policy %d", policy);
    if (min_cells_per_step < 0) return fail(BGK_EINVAL, "min_cells_per_step %lld", min_cells_per_step);
    s->policy = policy;
    s->min_cells = min_cells_per_step;
    return 0;
}

int bgks_destroy(bgks_server* s)
{
    if (!s) return 0;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& g : s->groups) {
        if (g.d_slots) (void)hipFree(g.d_slots);
        if (g.d_owner) (void)hipFree(g.d_owner);
        if (g.h_slots) (void)hipHostFree(g.h_slots);
        if (g.h_owner) (void)hipHostFree(g.h_owner);
    }
    if (s->arena_mem) (void)hipFree(s->arena_mem);
    if (s->scratch) (void)hipFree(s->scratch);
    if (s->hscratch) (void)hipHostFree(s->hscratch);
    if (s->d_flags) (void)hipFree(s->d_flags);
    if (s->h_flags) (void)hipHostFree(s->h_flags);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return 0;
}

int bgks_submit(bgks_server* s, const bgks_request* r, long long* id)
{
    if (!s || !r || !id) return fail(BGK_EINVAL, "null argument");
    if (r->steps < 0) return fail(BGK_EINVAL, "steps %lld", r->steps);
    if (r->rounding != 0 && r->rounding != 1) return fail(BGK_EINVAL, "rounding %d", r->rounding);
    auto in = std::make_unique<Instance>();
    if (int rc = handle_setup(in->h, request_config(*r, s->device))) return rc;
    in->h.borrowed = true;
    in->h.stream = s->stream;
    in->h.scratch = s->scratch;
    in->h.hscratch = s->hscratch;
    in->len = round_up((long)lattice_bytes(in->h), (long)kArenaAlign);
    if (in->len > s->arena.size)
        return fail(BGK_ENOMEM, "request needs %.3f GB, the budget is %.3f GB", in->len / 1e9, s->arena.size / 1e9);
    in->req = *r;
    in->id = s->next_id++;
    in->t_submit = clock_now();
    *id = in->id;
    s->queue.push_back(in->id);
    s->all[in->id] = std::move(in);
    return 0;
}

int bgks_advance(bgks_server* s, long long* done_ids, int max_done, int* n_done)
{
    if (!s) return fail(BGK_EINVAL, "null server");
    if (n_done) *n_done = 0;
    ENG_HIP(hipSetDevice(s->device));
    if (int rc = admit(*s)) return rc;

    std::vector<Instance*> active;
    for (auto& kv : s->all)
        if (kv.second->state == BGKS_ACTIVE) active.push_back(kv.second.get());
    if (s->policy == BGKS_SHORTEST_FIRST && !active.empty()) {
        std::vector<long long> rem(active.size()), cells(active.size());
        std::vector<int> skip(active.size());
        for (size_t q = 0; q < active.size(); ++q) {
            rem[q] = remaining_work(*active[q]);
            cells[q] = (long long)active[q]->h.cfg.l * active[q]->h.cfg.m * active[q]->h.cfg.n;
            skip[q] = active[q]->skipped;
        }
        const std::vector<int> sel =
            plan_shortest(rem.data(), cells.data(), skip.data(), (int)active.size(), s->min_cells, kMaxSkips);
        std::vector<char> chosen(active.size(), 0);
        for (int q : sel) chosen[q] = 1;
        std::vector<Instance*> run;
        for (size_t q = 0; q < active.size(); ++q) {
            if (chosen[q]) {
                active[q]->skipped = 0;
                run.push_back(active[q]);
            } else {
                ++active[q]->skipped;
            }
        }
        active.swap(run);
    }
    for (long long id : s->queue) ++s->all.at(id)->skipped;
    long long K = s->chunk;
    long long cells_in_flight = 0;
    for (Instance* in : active) {
        K = std::min(K, in->req.steps - in->h.steps);
        cells_in_flight += (long long)in->h.cfg.l * in->h.cfg.m * in->h.cfg.n;
    }
    // chunk-time target: as many steps as the last chunk's rate per cell
    // step fits into chunk_ms (at least one)
    if (s->chunk_ms > 0 && s->ns_per_cell_step > 0 && cells_in_flight > 0) {
        const double per_step_ms = s->ns_per_cell_step * (double)cells_in_flight * 1e-6;
        K = std::min(K, std::max(1LL, (long long)(s->chunk_ms / per_step_ms)));
    }
    K = std::max(K, 0LL);

    bool ran = false;
    if (K > 0 && !active.empty()) {
        std::vector<Instance*> members[kGroups];
        for (Instance* in : active)
            members[group_of(in->h.cfg.storage, in->h.cfg.rounding, in->h.compact)].push_back(in);
        std::vector<SubBatch> subs[kGroups];
        for (int g = 0; g < kGroups; ++g) {
            if (members[g].empty()) continue;
            int rc = by_group(g, [&](auto tag, auto) {
                return prepare_group<decltype(tag)>(*s, s->groups[g], members[g], subs[g], s->stream);
            });
            if (rc) return rc;
        }
        const double t0 = clock_now();
        // each sub-batch runs all K steps before the next one, so that
        // consecutive launches stay inside its lattices.  One stream: two
        // sub-batches streaming side by side on a second stream ran the mix
        // about 6 % slower than one after the other.
        for (int g = 0; g < kGroups; ++g)
            for (const SubBatch& sb : subs[g])
                for (long long t = 0; t < K; ++t) {
                    by_group(g, [&](auto tag, auto sr) {
                        launch_group<decltype(tag), decltype(sr)::value>(s->groups[g], sb, (int)t, s->stream,
                                                                          g >= 8);
                        return 0;
                    });
                    ++s->launches;
                }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(BGK_EHIP, "batched step launch: %s", hipGetErrorString(e));
        ENG_HIP(hipStreamSynchronize(s->stream));
        s->ns_per_cell_step = (clock_now() - t0) * 1e9 / ((double)K * (double)cells_in_flight);
        for (Instance* in : active) {
            in->h.steps += K;
            in->h.cur = (int)((in->h.cur + K) & 1);
            s->cell_steps += (long long)in->h.cfg.l * in->h.cfg.m * in->h.cfg.n * K;
        }
        ++s->chunks;
        ran = true;
        ENG_HIP(hipMemcpy(s->h_flags, s->d_flags, kMaxFlags * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    ENG_HIP(hipStreamSynchronize(s->stream));
    const double now = clock_now();

    int nd = 0;
    for (Instance* in : active) {
        // a non-finite value anywhere in the lattice retires the request now
        const bool bad = ran && s->h_flags[in->flag] != 0;
        if (!bad && in->h.steps < in->req.steps) continue;
        if (int rc = bgk_diagno(&in->h, in->diag)) return rc;
        in->unstable = bad ? 1 : 0;
        for (double v : in->diag)
            if (!std::isfinite(v)) in->unstable = 1;
        in->state = BGKS_DONE;
        in->t_done = now;
        if (!in->req.keep) free_lattice(*s, *in);
        else if (in->flag >= 0) {
            s->free_flags.push_back(in->flag);
            in->flag = -1;
        }
        if (done_ids && nd < max_done) done_ids[nd] = in->id;
        ++nd;
    }
    if (n_done) *n_done = nd;
    // requests that fit the memory just released start with the next chunk
    return admit(*s);
}

int bgks_status(bgks_server* s, long long id, int* state, long long* steps_done)
{
    if (!s) return fail(BGK_EINVAL, "null server");
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(BGK_EINVAL, "unknown request %lld", id);
    if (state) *state = it->second->state;
    if (steps_done) *steps_done = it->second->h.steps;
    return 0;
}

int bgks_result(bgks_server* s, long long id, double diag[5], double times[3], int* unstable)
{
    if (!s) return fail(BGK_EINVAL, "null server");
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(BGK_EINVAL, "unknown request %lld", id);
    const Instance& in = *it->second;
    if (in.state != BGKS_DONE) return fail(BGK_EBUSY, "request %lld has not finished", id);
    if (diag) std::memcpy(diag, in.diag, sizeof in.diag);
    if (times) {
        times[0] = in.t_submit;
        times[1] = in.t_admit;
        times[2] = in.t_done;
    }
    if (unstable) *unstable = in.unstable;
    return 0;
}

int bgks_get_population(bgks_server* s, long long id, int d, double* out)
{
    if (!s) return fail(BGK_EINVAL, "null server");
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(BGK_EINVAL, "unknown request %lld", id);
    Instance& in = *it->second;
    if (!in.has_mem) return fail(BGK_EINVAL, "request %lld holds no lattice (not admitted, or released)", id);
    ENG_HIP(hipStreamSynchronize(s->stream));
    return bgk_get_population(&in.h, d, out);
}

int bgks_release(bgks_server* s, long long id)
{
    if (!s) return fail(BGK_EINVAL, "null server");
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(BGK_EINVAL, "unknown request %lld", id);
    Instance& in = *it->second;
    if (in.state == BGKS_ACTIVE) return fail(BGK_EBUSY, "request %lld is running", id);
    if (in.state == BGKS_QUEUED) s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
    free_lattice(*s, in);
    s->all.erase(it);
    return 0;
}

int bgks_stats(bgks_server* s, long long out[8])
{
    if (!s || !out) return fail(BGK_EINVAL, "null argument");
    long long q = 0, a = 0, d = 0;
    for (auto& kv : s->all) {
        q += kv.second->state == BGKS_QUEUED;
        a += kv.second->state == BGKS_ACTIVE;
        d += kv.second->state == BGKS_DONE;
    }
    out[0] = q;
    out[1] = a;
    out[2] = d;
    out[3] = (long long)s->arena.used();
    out[4] = (long long)s->arena.size;
    out[5] = s->launches;
    out[6] = s->cell_steps;
    out[7] = s->chunks;
    return 0;
}

}  // extern "C"
