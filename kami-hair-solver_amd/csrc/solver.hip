// Host control of the hair step on one MI355X: device buffers, LBVH
// rebuild, the Newton / line-search loop and the adaptive frame loop
// (reference src/cuda_backend.cu:341-723, src/solver.cpp:449-893).
#include "solver.hpp"

#include "model_kernels.hpp"
#include "root_kernels.hpp"
#include "step_kernels.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>

#include <unistd.h>

namespace khs {

// A kernel launch on a solver's stream, counted (khsxGetLaunchCounts).
#define KHS_LAUNCH(state, ...)                                                                                      \
    do {                                                                                                          \
        ++(state).launch_count;                                                                                   \
        hipLaunchKernelGGL(__VA_ARGS__);                                                                          \
    } while (0)

#define HIP_TRY(call)                                                                                         \
    do {                                                                                                      \
        const hipError_t khs_err_ = (call);                                                                   \
        if (khs_err_ != hipSuccess)                                                                           \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(khs_err_) + " in " #call); \
    } while (0)

namespace {

constexpr uint64_t kCheckpointMagic = 0x3154504348534bULL;
constexpr double kPi = 3.1415926535897932384626433832795;
const double kInf = std::numeric_limits<double>::infinity();
// LDS the block solve reserves without using it, so that at most five of
// its one-wavefront workgroups share a CU: the longest strands, dispatched
// first, then run nearly alone on their SIMDs (measured 8% faster than
// letting two wavefronts share each SIMD).
constexpr unsigned kSolveLdsReserve = 28 * 1024;
// Blocks of the swept step-bound kernel: its gathering blocks, then the
// list blocks.
constexpr unsigned kSweptBoundBlocks = khs::dev::kSweptGatherBlocks + khs::kPairBlocks;

template <class T>
class Buffer {
public:
    Buffer() = default;
    Buffer(const Buffer &) = delete;
    Buffer &operator=(const Buffer &) = delete;
    ~Buffer() { reset(); }
    void allocate(size_t count)
    {
        reset();
        count_ = count;
        if (count) HIP_TRY(hipMalloc(&ptr_, count * sizeof(T)));
        owned_ = true;
    }
    // `count` elements of another buffer's allocation, which outlives this
    void view(T *ptr, size_t count)
    {
        reset();
        ptr_ = ptr;
        count_ = count;
        owned_ = false;
    }
    // Exchanges the two buffers' memory (no copy).
    void swap(Buffer &o)
    {
        std::swap(ptr_, o.ptr_);
        std::swap(count_, o.count_);
        std::swap(owned_, o.owned_);
    }
    void reset()
    {
        if (ptr_ && owned_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        count_ = 0;
    }
    void upload(const T *src, size_t count, size_t offset = 0)
    {
        if (offset + count > count_) throw std::runtime_error("device buffer overflow");
        if (count) HIP_TRY(hipMemcpy(ptr_ + offset, src, count * sizeof(T), hipMemcpyHostToDevice));
    }
    void download(T *dst, size_t count, size_t offset = 0) const
    {
        if (offset + count > count_) throw std::runtime_error("device buffer overflow");
        if (count) HIP_TRY(hipMemcpy(dst, ptr_ + offset, count * sizeof(T), hipMemcpyDeviceToHost));
    }
    T *get() const { return ptr_; }
    size_t size() const { return count_; }
    size_t bytes() const { return count_ * sizeof(T); }

private:
    T *ptr_ = nullptr;
    size_t count_ = 0;
    bool owned_ = true;
};

unsigned grid_for(size_t n, unsigned block) { return unsigned((n + block - 1) / block); }

std::vector<double> flatten(const std::vector<Vec3> &v)
{
    std::vector<double> out(v.size() * 3);
    for (size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
    return out;
}

std::vector<Vec3> to_vec3(const KhsVec3 *p, uint32_t n)
{
    std::vector<Vec3> out(n);
    for (uint32_t i = 0; i < n; ++i) out[i] = {p[i].x, p[i].y, p[i].z};
    return out;
}

// Staging of the allocation-free broad-phase query
// (khsxBroadphasePairsNoAlloc): a pair counter on its own 256-byte line,
// then at most kNoAllocPairs (element, triangle) pairs, 8 MiB in all. Each
// GPU holds its own copy until the process ends. Host code reaches it only
// through the address hipGetSymbolAddress returns for the current device
// (no_alloc_staging), never through the array's name.
constexpr uint64_t kNoAllocPairs = KHSX_NO_ALLOC_PAIRS;
constexpr uint64_t kNoAllocHeader = 64;  // uint32 words before the pairs
static __device__ __align__(256) uint32_t scratch[kNoAllocHeader + 2 * kNoAllocPairs];

constexpr int kMaxDevices = 64;
std::mutex g_staging_mutex;
uint32_t *g_staging[kMaxDevices] = {};

// Looks up the staging array's device address once per device, with that
// device current; later calls return the kept address.
void prepare_no_alloc_staging()
{
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("HIP device ordinal out of range.");
    std::lock_guard<std::mutex> lock(g_staging_mutex);
    if (g_staging[device]) return;
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw std::runtime_error("No device address for the broad-phase staging array.");
    g_staging[device] = static_cast<uint32_t *>(ptr);
}

uint32_t *no_alloc_staging()
{
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    std::lock_guard<std::mutex> lock(g_staging_mutex);
    if (device < 0 || device >= kMaxDevices || !g_staging[device])
        throw std::runtime_error("The broad-phase staging array was not set up on this device.");
    return g_staging[device];
}

}  // namespace

struct DeviceState {
    uint64_t launch_count = 0;  // kernel launches and device-to-device copies on `stream`
    uint64_t wait_count = 0;    // host waits for `stream`
    hipStream_t stream = nullptr;
    hipEvent_t frame_begin = nullptr, frame_end = nullptr;
    uint32_t N = 0, S = 0, V = 0, T = 0, P = 0;
    // rod model
    Buffer<double> frame, rest, mass, rmass;
    Buffer<uint32_t> flags, strand_of, strand_first, strand_count, strand_fixed, mapping;
    Buffer<uint32_t> solve_order;  // strands by decreasing free node count, for the solve's dispatch
    Buffer<uint32_t> root_elements;  // contact elements with a fixed left and a free right node
    uint32_t root_element_count = 0;
    Buffer<unsigned long long> ccd_bound;  // two alternating launch-wide step bounds of the list CCD
    uint32_t ccd_slot = 0;
    Buffer<uint32_t> bind_left, bind_right;  // per node: the original points its position lerps
    Buffer<double> bind_t, root_points;      // lerp weight; staging for one frame of original points
    Buffer<unsigned> root_flag;              // set when a frame's targets have a degenerate tangent
    Buffer<unsigned> solve_failed;  // two alternating failure flags of the block solve
    uint32_t failed_slot = 0;
    dev::TrialStep trial_step{};     // how the next trial's first kernel sets alpha
    // state and solver vectors
    Buffer<double> q, q_old, vel, q_trial, snap_q, snap_v, grad, dir, precond, Dblk, Ublk, Cp, dp, dv;
    // D, U and the gradient in one allocation (D | U | g), which the fused
    // block solve reads through one buffer resource; Dblk, Ublk and grad
    // are views of it
    Buffer<double> solve_in;
    // root targets (3 per node)
    Buffer<double> root_cur, rot_cur, root_pend, rot_pend, root_anim, rot_anim;
    Buffer<double> ring_q, ring_v;  // rollback ring: 6N doubles per slot each
    // collider (3 per vertex)
    Buffer<double> col_cur, col_old, col_start, col_pend, col_anim;
    Buffer<uint32_t> tris;
    // LBVH
    Buffer<double> tri_box, tri_box_static, bound_partials, bounds;
    Buffer<uint32_t> keys, keys_sorted, ids, ids_sorted;
    Buffer<BvhNode> bvh;
    Buffer<int> parent_internal, parent_leaf;
    Buffer<unsigned> counters, refit_ready_count;
    Buffer<uint32_t> refit_ready;  // chunk-spanning nodes whose children are done
    Buffer<unsigned char> sort_temp;
    size_t sort_temp_bytes = 0;
    // per-element contact candidate lists (valid for the current substep)
    Buffer<uint32_t> cache_count, cache_tris, cache_listed, cache_list, cache_list_count, cache_pending,
        cache_pending_count, cache_pending_epoch, cache_snapshot;
    uint32_t pending_slot = 0;  // which pending counter the next swept pass uses
    uint32_t swept_epoch = 0;   // swept passes so far (marks of the elements each one queues)
    Buffer<double> cache_box;
    double cache_skin = 0.0;
    bool cache_valid = false;
    // best_tri and list_words hold the list pass at the positions q now has:
    // the accepted line-search trial's, whose check and list pass ran at
    // alpha d + base, the expression the trial evaluation wrote into q_trial
    // (now q). The
    // gradient evaluation at that iterate then needs neither again.
    bool closest_at_q = false;
    // Every contact list covers the swept boxes of q to q + dir (set by the
    // swept pass, cleared when dir or q change otherwise).
    bool swept_covers = false;
    Buffer<uint32_t> best_tri;  // closest triangle per element from the list pass
    Buffer<double> best_pair;   // and its closest pair (dev::kPairWords per element)
    // reductions and flags
    Buffer<Partial> partials, reduced;
    Buffer<unsigned> flag;
    Buffer<unsigned long long> key64;
    Buffer<SweepRecord> sweep;
    Buffer<double> out3;
    Buffer<uint32_t> pairs;
    Buffer<double> alpha;  // line-search step read by the trial kernels
    Buffer<unsigned long long> list_words;  // per-block candidate counts of the narrow-phase list pass
    Buffer<double> list_alpha;              // per-block smallest step of the CCD list pass
    // per-strand materials: (E, r) per node of its strand, and per strand
    // the mass and rotational mass per length; empty without them
    Buffer<double> elem_material, strand_density;
    DirectionChoice *choice() const { return reinterpret_cast<DirectionChoice *>(reduced.get() + 3); }

    // Event timing of kernel classes (enabled on request) and of substeps.
    struct Pending {
        hipEvent_t begin, end;
        double *ms;
        uint64_t *launches;
    };
    bool timing = false;
    std::vector<hipEvent_t> event_pool;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> substep_events;
    std::vector<double> substep_ms;
    KhsxKernelTiming totals{};
    hipEvent_t take_event()
    {
        if (event_pool.empty()) {
            // timing only: no system-scope fence (no L2 writeback) at record
            hipEvent_t e;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            return e;
        }
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t tic()
    {
        if (!timing) return nullptr;
        hipEvent_t e = take_event();
        HIP_TRY(hipEventRecord(e, stream));
        return e;
    }
    void toc(hipEvent_t begin, double &ms, uint64_t &launches)
    {
        if (!timing || !begin) return;
        hipEvent_t e = take_event();
        HIP_TRY(hipEventRecord(e, stream));
        pending.push_back({begin, e, &ms, &launches});
    }
    // Called after the stream has been synchronized.
    void settle()
    {
        for (const Pending &p : pending) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, p.begin, p.end));
            *p.ms += ms;
            *p.launches += 1;
            event_pool.push_back(p.begin);
            event_pool.push_back(p.end);
        }
        pending.clear();
    }

    ~DeviceState()
    {
        for (const Pending &p : pending) {
            (void)hipEventDestroy(p.begin);
            (void)hipEventDestroy(p.end);
        }
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
        for (hipEvent_t e : substep_events) (void)hipEventDestroy(e);
        if (host_view) (void)hipHostFree(host_view);
        if (cache_host) (void)hipHostFree(cache_host);
        if (frame_begin) (void)hipEventDestroy(frame_begin);
        if (frame_end) (void)hipEventDestroy(frame_end);
        if (stream) (void)hipStreamDestroy(stream);
    }

    RodView rods() const
    {
        return RodView{N,           S,           frame.get(),        rest.get(),         mass.get(),
                       rmass.get(), flags.get(), strand_of.get(), strand_first.get(), strand_count.get(),
                       strand_fixed.get()};
    }
    // Queries test leaves against the boxes at the current collider position
    // (the tree may be refit over the substep's motion); the sweep test
    // alone uses the tree's own, swept, leaf boxes.
    ColliderView collider() const
    {
        return ColliderView{V, T, col_cur.get(), col_old.get(), tris.get(), bvh.get(), tri_box_static.get()};
    }
    ColliderView swept_collider() const
    {
        return ColliderView{V, T, col_cur.get(), col_old.get(), tris.get(), bvh.get(), nullptr};
    }
    ContactCache cache() const
    {
        return ContactCache{cache_skin,       cache_count.get(),      cache_tris.get(),        cache_box.get(),
                            cache_listed.get(), cache_list.get(),  cache_list_count.get(), cache_pending.get(),
                            cache_pending_count.get() + pending_slot, tri_box_static.get(),
                            cache_pending_epoch.get(), swept_epoch, cache_snapshot.get()};
    }

    size_t resident_bytes() const
    {
        size_t b = 0;
        for (const Buffer<double> *x : {&frame, &rest, &mass, &rmass, &q, &q_old, &vel, &q_trial, &snap_q, &snap_v,
                                        &grad, &dir, &precond, &Dblk, &Ublk, &Cp, &dp, &dv, &root_cur, &rot_cur,
                                        &root_pend, &rot_pend, &root_anim, &rot_anim, &ring_q, &ring_v, &col_cur,
                                        &col_old, &col_start,
                                        &col_pend, &col_anim, &tri_box, &tri_box_static, &bound_partials, &bounds, &out3})
            b += x->bytes();
        for (const Buffer<uint32_t> *x : {&flags, &strand_of, &strand_first, &strand_count, &strand_fixed, &mapping,
                                          &tris, &keys, &keys_sorted, &ids, &ids_sorted})
            b += x->bytes();
        b += bvh.bytes() + parent_internal.bytes() + parent_leaf.bytes() + counters.bytes() + sort_temp.bytes() +
             partials.bytes() + reduced.bytes() + cache_count.bytes() + cache_tris.bytes() + cache_box.bytes() +
             cache_listed.bytes() + cache_list.bytes() + cache_list_count.bytes() + cache_pending.bytes() +
             cache_pending_epoch.bytes() + cache_snapshot.bytes() +
             best_tri.bytes() + best_pair.bytes();
        return b;
    }

    Partial reduce(uint32_t blocks)
    {
        reduce_async(blocks, kEvalSlot);
        return read_back().eval;
    }
    // Folds the partials into reduced[slot] without waiting (and applies
    // the convergence test when one is given).
    void reduce_async(uint32_t blocks, int slot, dev::ConvergenceTest test = {0.0, 0.0, 0.0, nullptr},
                      dev::ListPartials lists = {nullptr, nullptr, 0})
    {
        KHS_LAUNCH(*this, dev::k_reduce, dim3(1), dim3(dev::kReduceBlock), 0, stream, partials.get(), blocks,
                           reduced.get() + slot, mirror_slot(slot), test, lists);
        HIP_TRY(hipGetLastError());
    }
    // One wait for everything the host reads after a batch of launches.
    struct HostView {
        Partial eval, steep, ccd;
        union {
            DirectionChoice choice;
            Partial choice_space;
        };
        Partial grad, dir;
    };
    // The reductions and choice kernels write host_view as they run (a
    // mirror of `reduced` in coherent pinned memory), so a read-back is a
    // wait and no copy.
    HostView read_back()
    {
        static_assert(sizeof(HostView) == 6 * sizeof(Partial), "one Partial per slot of reduced");
        wait();
        return *host_view;
    }
    Partial *mirror_slot(int slot) const { return reinterpret_cast<Partial *>(host_view_device) + slot; }
    DirectionChoice *mirror_choice() const
    {
        return reinterpret_cast<DirectionChoice *>(reinterpret_cast<Partial *>(host_view_device) + 3);
    }
    void wait()
    {
        ++wait_count;
        HIP_TRY(hipStreamSynchronize(stream));
        settle();
    }
    // slots of reduced (3 holds the direction choice)
    static constexpr int kEvalSlot = 0, kSteepSlot = 1, kCcdSlot = 2, kGradSlot = 4, kDirSlot = 5;
    Buffer<unsigned> stop;  // set by the gradient's convergence test: work enqueued ahead skips itself
    bool skipping = false;  // inside the Newton loop, where that skip applies
    const unsigned *skip() const { return skipping ? stop.get() : nullptr; }
    HostView *host_view = nullptr;  // pinned, coherent
    void *host_view_device = nullptr;  // its device address
    double *cache_host = nullptr;      // pinned staging of one cache frame (3P doubles)

    uint32_t tiled_blocks() const { return grid_for(size_t(N) + 1, kNodesPerBlock); }
    uint32_t eval_blocks() const { return grid_for(size_t(N) + 1, kEvalNodesPerBlock); }
    uint32_t solve_blocks() const { return grid_for(S, dev::kSolveGroups); }
    uint32_t twisted_blocks() const { return grid_for(S, dev::kTwistGroups); }
    uint32_t node_blocks() const { return grid_for(N, kBlock); }
};

struct Solver::Reduced {
    double objective, elastic, contact, kinetic, grad_norm2, dir_norm2, min_gap;
    unsigned long long candidates, active;
    bool feasible;
};

Solver::Solver(const KhsSolverDesc &desc) : desc_(desc)
{
    build_stats_.struct_size = sizeof(KhsBuildStats);
    step_stats_.struct_size = sizeof(KhsStepStats);
    step_stats_.phase = KHS_PHASE_IDLE;
    gpu_stats_.struct_size = sizeof(KhsGpuStats);
    clear_failure();
}

Solver::~Solver() = default;

void Solver::release_device() { dev_.reset(); }

KhsResult Solver::fail(KhsResult result, const std::string &message)
{
    error_ = message;
    step_stats_.phase = KHS_PHASE_FAILED;
    phase_.store(KHS_PHASE_FAILED);
    return result;
}

void Solver::clear_failure()
{
    failure_ = {};
    failure_.struct_size = sizeof(failure_);
    failure_.strand_index = 0xffffffffu;
    failure_.element_index = 0xffffffffu;
    failure_.collider_triangle_index = 0xffffffffu;
}

void Solver::set_failure_context(uint32_t kind, uint32_t frame, uint32_t sub, uint32_t attempted, uint32_t maximum,
                                 uint32_t attempts)
{
    failure_.kind = kind;
    failure_.frame_index = frame;
    failure_.substep = sub;
    failure_.requested_substeps = desc_.substeps;
    failure_.attempted_substeps = attempted;
    failure_.maximum_substeps = maximum;
    failure_.adaptive_attempt_count = attempts;
}

// ------------------------------------------------------------------ input
KhsResult Solver::set_hair(const KhsVec3 *points, uint32_t point_count, const uint32_t *offsets, uint32_t strand_count,
                           const KhsHairMaterial &m)
{
    // reference solver.cpp:136-177
    if (!points || !offsets || point_count < 2 || strand_count == 0 || offsets[0] != 0 ||
        offsets[strand_count] != point_count) {
        error_ = "Invalid hair points or strand offsets.";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    for (uint32_t s = 0; s < strand_count; ++s)
        if (offsets[s + 1] <= offsets[s] || offsets[s + 1] > point_count) {
            error_ = "Hair strand offsets are not increasing.";
            return KHS_ERROR_INVALID_HAIR;
        }
    if (!(m.density > 0.0 && m.radius > 0.0 && m.young_modulus > 0.0 && m.poisson_ratio > -1.0 &&
          m.poisson_ratio < 0.5 && m.shear_correction > 0.0 && m.contact_stiffness > 0.0 && m.barrier_distance > 0.0 &&
          m.friction >= 0.0 && m.friction_smoothing > 0.0)) {
        error_ = "Hair material is out of range.";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    std::vector<Vec3> pts = to_vec3(points, point_count);
    for (uint32_t i = 0; i < point_count; ++i)
        if (!is_finite(pts[i])) {
            error_ = "Hair point " + std::to_string(i) + " is not finite.";
            return KHS_ERROR_INVALID_HAIR;
        }
    built_ = false;
    animation_ready_ = false;
    ring_frame_.clear();
    release_device();
    points_ = std::move(pts);
    offsets_.assign(offsets, offsets + strand_count + 1);
    strand_E_.clear();
    strand_r_.clear();
    material_ = m;
    has_hair_ = true;
    error_.clear();
    return KHS_OK;
}

KhsResult Solver::set_collider(const KhsVec3 *vertices, uint32_t vertex_count, const KhsTriangle *triangles,
                               uint32_t triangle_count)
{
    // reference solver.cpp:179-225
    if (triangle_count == 0) {
        has_collider_ = false;
        collider_ = ColliderMesh{};
        built_ = false;
        return KHS_OK;
    }
    if (!vertices || !triangles || vertex_count < 3) {
        error_ = "Invalid collider vertices or triangles.";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    ColliderMesh mesh;
    const KhsResult r = filter_collider(vertices, vertex_count, triangles, triangle_count, mesh, error_);
    if (r != KHS_OK) return r;
    collider_ = std::move(mesh);
    has_collider_ = true;
    built_ = false;
    error_.clear();
    return KHS_OK;
}

// ------------------------------------------------------------------ build
KhsResult Solver::build()
{
    // reference solver.cpp:449-556
    if (!has_hair_) {
        error_ = "Set hair curves first.";
        return KHS_ERROR_INVALID_STATE;
    }
    built_ = false;
    animation_ready_ = false;
    ring_frame_.clear();
    release_device();
    step_stats_.phase = KHS_PHASE_PREPARING;
    if (has_collider_) validate_collider(collider_);
    KhsResult r = upload_model();
    if (r != KHS_OK) {
        step_stats_.phase = KHS_PHASE_FAILED;
        release_device();
        return r;
    }
    if (!strand_r_.empty()) apply_strand_materials();
    double initial_gap = kInf;
    if (has_collider_) {
        r = initial_feasibility();
        if (r != KHS_OK) {
            step_stats_.phase = KHS_PHASE_FAILED;
            release_device();
            return r;
        }
        initial_gap = build_stats_.initial_minimum_gap;
    }
    built_ = true;
    build_stats_ = {};
    build_stats_.struct_size = sizeof(KhsBuildStats);
    build_stats_.strand_count = model_.strand_count;
    build_stats_.original_point_count = uint32_t(points_.size());
    build_stats_.internal_node_count = model_.node_count;
    build_stats_.element_count = model_.element_count;
    build_stats_.fixed_node_count = model_.fixed_nodes;
    build_stats_.collider_vertex_count = uint32_t(collider_.vertices.size());
    build_stats_.collider_triangle_count = collider_.triangle_count();
    build_stats_.excluded_collider_triangle_count = collider_.excluded;
    build_stats_.collider_boundary_edge_count = collider_.boundary_edges;
    build_stats_.collider_nonmanifold_edge_count = collider_.nonmanifold_edges;
    build_stats_.collider_inconsistent_edge_count = collider_.inconsistent_edges;
    build_stats_.collider_inverted_closed_component_count = collider_.inverted;
    build_stats_.merged_zero_length_segment_count = model_.merged_segments;
    build_stats_.virtual_extension_strand_count = model_.extension_strands;
    build_stats_.virtual_extension_node_count = model_.extension_nodes;
    build_stats_.virtual_extension_rest_length = model_.extension_length;
    build_stats_.degree_of_freedom_count = model_.free_dofs;
    build_stats_.estimated_bytes = dev_->resident_bytes();
    build_stats_.initial_minimum_gap = initial_gap;
    gpu_stats_ = {};
    gpu_stats_.struct_size = sizeof(KhsGpuStats);
    gpu_stats_.resident_bytes = dev_->resident_bytes();
    step_stats_ = {};
    step_stats_.struct_size = sizeof(KhsStepStats);
    step_stats_.phase = KHS_PHASE_IDLE;
    phase_.store(KHS_PHASE_IDLE);
    error_.clear();
    return KHS_OK;
}

// The rod model on the device (reference solver.cpp:227-373): a counting
// pass per strand, the node layout scanned on the host from the per-strand
// counts (S values), then a pass per strand that writes nodes, bindings,
// frames, rest strains and masses into the buffers the step reads.
KhsResult Solver::build_model_on_device()
{
    DeviceState &d = *dev_;
    RodModel &m = model_;
    m = RodModel{};
    const uint32_t S = uint32_t(offsets_.size() - 1);
    d.S = S;
    d.P = uint32_t(points_.size());
    const std::vector<double> flat = flatten(points_);
    d.root_points.allocate(flat.size());
    d.root_points.upload(flat.data(), flat.size());
    Buffer<uint32_t> offsets;
    offsets.allocate(S + 1);
    offsets.upload(offsets_.data(), S + 1);
    Buffer<dev::StrandShape> shape;
    shape.allocate(S);
    const double hmax = desc_.maximum_element_length, minimum = desc_.minimum_dynamic_length;
    KHS_LAUNCH(d, dev::k_model_count, dim3(grid_for(S, kBlock)), dim3(kBlock), 0, d.stream, S,
                       (const double *)d.root_points.get(), (const uint32_t *)offsets.get(), hmax, minimum,
                       shape.get());
    HIP_TRY(hipGetLastError());
    std::vector<dev::StrandShape> sh(S);
    HIP_TRY(hipMemcpyAsync(sh.data(), shape.get(), shape.bytes(), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));

    m.strand_count = S;
    m.strand_first.resize(S);
    m.strand_count_of.resize(S);
    m.strand_fixed.resize(S);
    m.strand_real.resize(S);
    uint64_t total = 0;
    for (uint32_t s = 0; s < S; ++s) {
        const dev::StrandShape &x = sh[s];
        if (x.degenerate) {
            error_ = "Strand " + std::to_string(s) + " has all its points at one position.";
            return KHS_ERROR_INVALID_HAIR;
        }
        m.strand_first[s] = uint32_t(std::min<uint64_t>(total, UINT32_MAX));
        m.strand_count_of[s] = x.nodes;
        m.strand_real[s] = x.real_nodes;
        m.strand_fixed[s] = std::min<uint32_t>(desc_.fixed_root_nodes, x.nodes);
        m.merged_segments += x.merged;
        if (x.extension_nodes) {
            ++m.extension_strands;
            m.extension_nodes += x.extension_nodes;
            m.extension_length += x.extension_length;
        }
        m.fixed_nodes += m.strand_fixed[s];
        m.free_dofs += 6 * uint64_t(x.nodes - m.strand_fixed[s]);
        total += x.nodes;
    }
    if (total > dev::kMaxSolveNodes) {
        error_ = "The model has " + std::to_string(total) + " nodes; the block solve addresses at most " +
                 std::to_string(dev::kMaxSolveNodes) + " per solver (32-bit byte offsets into the Hessian blocks).";
        return KHS_ERROR_OUT_OF_MEMORY;
    }
    m.node_count = uint32_t(total);
    m.element_count = m.node_count - S;
    d.N = m.node_count;
    const size_t N = d.N;

    d.q.allocate(6 * N);
    d.frame.allocate(9 * N);
    d.rest.allocate(8 * N);
    d.mass.allocate(N);
    d.rmass.allocate(N);
    d.flags.allocate(N);
    d.strand_of.allocate(N);
    d.bind_left.allocate(N);
    d.bind_right.allocate(N);
    d.bind_t.allocate(N);
    d.mapping.allocate(d.P);
    d.root_flag.allocate(1);
    d.strand_first.allocate(S);
    d.strand_first.upload(m.strand_first.data(), S);
    d.strand_count.allocate(S);
    d.strand_count.upload(m.strand_count_of.data(), S);
    d.strand_fixed.allocate(S);
    d.strand_fixed.upload(m.strand_fixed.data(), S);
    Buffer<double> strand_max;
    strand_max.allocate(S);
    const MassDensity rho = mass_density(material_);
    dev::ModelOut o{d.q.get(),        d.frame.get(),    d.rest.get(),       d.mass.get(),   d.rmass.get(),
                    d.flags.get(),    d.strand_of.get(), d.bind_left.get(), d.bind_right.get(), d.bind_t.get(),
                    d.mapping.get(),  strand_max.get()};
    KHS_LAUNCH(d, dev::k_model_nodes, dim3(grid_for(S, kBlock)), dim3(kBlock), 0, d.stream, S,
                       (const double *)d.root_points.get(), (const uint32_t *)offsets.get(), hmax, minimum,
                       desc_.fixed_root_nodes, (const uint32_t *)d.strand_first.get(), rho.translational,
                       rho.rotational, o);
    HIP_TRY(hipGetLastError());
    std::vector<double> maxima(S);
    HIP_TRY(hipMemcpyAsync(maxima.data(), strand_max.get(), strand_max.bytes(), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    for (double v : maxima) m.maximum_rest_length = std::max(m.maximum_rest_length, v);
    return KHS_OK;
}

KhsResult Solver::upload_model()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        error_ = "No HIP GPU is available.";
        return KHS_ERROR_INVALID_STATE;
    }
    HIP_TRY(hipSetDevice(0));
    prepare_no_alloc_staging();
    dev_ = std::make_unique<DeviceState>();
    DeviceState &d = *dev_;
    HIP_TRY(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&d.frame_begin));
    HIP_TRY(hipEventCreate(&d.frame_end));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&d.host_view), sizeof(DeviceState::HostView),
                          hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(d.host_view, 0, sizeof(DeviceState::HostView));
    HIP_TRY(hipHostGetDevicePointer(&d.host_view_device, d.host_view, 0));
    d.alpha.allocate(1);
    KhsResult built = build_model_on_device();
    if (built != KHS_OK) return built;
    const RodModel &m = model_;
    const size_t N = d.N;
    {
        // The solve is sequential along a strand: dispatching the longest
        // strands first keeps them off the tail of the launch.
        std::vector<uint32_t> order(d.S);
        for (uint32_t k = 0; k < d.S; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            return m.strand_count_of[x] - m.strand_fixed[x] > m.strand_count_of[y] - m.strand_fixed[y];
        });
        d.solve_order.allocate(d.S);
        d.solve_order.upload(order.data(), d.S);
    }
    {
        // Friction reads the closest triangle of these; the energy does not.
        // Per strand it is the element of the last fixed node, when the node
        // after it is free and not virtual.
        std::vector<uint32_t> roots;
        for (uint32_t s = 0; s < d.S; ++s) {
            const uint32_t fixed = m.strand_fixed[s];
            if (fixed >= 1 && fixed < m.strand_count_of[s] && fixed < m.strand_real[s])
                roots.push_back(m.strand_first[s] + fixed - 1);
        }
        d.root_element_count = uint32_t(roots.size());
        const double one = 1.0;
        unsigned long long bits[2];
        std::memcpy(&bits[0], &one, sizeof(one));
        bits[1] = bits[0];
        d.ccd_bound.allocate(2);
        d.ccd_bound.upload(bits, 2);
        const unsigned zeros[2] = {0u, 0u};
        d.solve_failed.allocate(2);
        d.solve_failed.upload(zeros, 2);
        d.root_elements.allocate(std::max<size_t>(1, roots.size()));
        if (!roots.empty()) d.root_elements.upload(roots.data(), roots.size());
    }

    // q = (position, zero rotation) from the model build, v = 0 (reference
    // cuda_backend.cu:360-371)
    const std::vector<double> zeros(6 * N, 0.0);
    for (Buffer<double> *b : {&d.q_old, &d.vel, &d.q_trial, &d.snap_q, &d.snap_v, &d.dir, &d.precond, &d.dp})
        b->allocate(6 * N);
    d.solve_in.allocate(78 * N);
    d.Dblk.view(d.solve_in.get(), 36 * N);
    d.Ublk.view(d.solve_in.get() + 36 * N, 36 * N);
    d.grad.view(d.solve_in.get() + 72 * N, 6 * N);
    ++d.launch_count;
    HIP_TRY(hipMemcpyAsync(d.q_old.get(), d.q.get(), d.q.bytes(), hipMemcpyDeviceToDevice, d.stream));
    ++d.launch_count;
    HIP_TRY(hipMemcpyAsync(d.snap_q.get(), d.q.get(), d.q.bytes(), hipMemcpyDeviceToDevice, d.stream));
    d.vel.upload(zeros.data(), zeros.size());
    d.snap_v.upload(zeros.data(), zeros.size());
    d.dir.upload(zeros.data(), zeros.size());
    d.grad.upload(zeros.data(), zeros.size());
    d.dv.allocate(3 * N);
    d.Cp.allocate(36 * N);
    // The root targets of the build's own points must have a frame
    // (reference solver.cpp:463-467); they are computed into the pending
    // targets as scratch. The targets then start at the node positions and
    // zero rotations (cuda_backend.cu:341-407).
    const std::vector<double> zero3(3 * N, 0.0);
    for (Buffer<double> *b : {&d.root_cur, &d.rot_cur, &d.root_pend, &d.rot_pend}) b->allocate(3 * N);
    if (!device_root_targets(points_, d.root_pend.get(), d.rot_pend.get())) {
        error_ = "Cannot build the initial root orientation.";
        return KHS_ERROR_INVALID_HAIR;
    }
    KHS_LAUNCH(d, dev::k_gather_positions, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.q.get(),
                       (const uint32_t *)nullptr, d.N, d.root_cur.get());
    HIP_TRY(hipGetLastError());
    ++d.launch_count;
    HIP_TRY(hipMemcpyAsync(d.root_pend.get(), d.root_cur.get(), d.root_cur.bytes(), hipMemcpyDeviceToDevice, d.stream));
    d.rot_cur.upload(zero3.data(), zero3.size());
    d.rot_pend.upload(zero3.data(), zero3.size());
    d.out3.allocate(3 * std::max<size_t>(N, d.P));
    const uint32_t max_blocks =
        std::max({d.tiled_blocks(), d.eval_blocks(), d.node_blocks(), d.solve_blocks(), d.twisted_blocks()}) + 1;
    d.partials.allocate(max_blocks + 16);
    d.list_words.allocate(kPairBlocks);
    d.list_alpha.allocate(kSweptBoundBlocks);
    d.reduced.allocate(6);  // reductions, with the direction choice in slot 3 (at most one Partial)
    {
        const unsigned zero = 0;
        d.stop.allocate(1);
        d.stop.upload(&zero, 1);
    }
    static_assert(sizeof(DirectionChoice) <= sizeof(Partial), "choice slot");
    d.flag.allocate(1);
    d.key64.allocate(1);
    d.sweep.allocate(1);

    if (has_collider_) {
        d.V = uint32_t(collider_.vertices.size());
        d.T = collider_.triangle_count();
        const std::vector<double> cv = flatten(collider_.vertices);
        for (Buffer<double> *b : {&d.col_cur, &d.col_old, &d.col_start, &d.col_pend}) {
            b->allocate(cv.size());
            b->upload(cv.data(), cv.size());
        }
        collider_pending_host_ = collider_.vertices;
        d.tris.allocate(collider_.triangles.size());
        d.tris.upload(collider_.triangles.data(), collider_.triangles.size());
        const size_t T = d.T;
        d.tri_box.allocate(6 * T);
        d.tri_box_static.allocate(kTriRecord * T);  // the triangle records
        d.bound_partials.allocate(6 * size_t(grid_for(T, kBlock)));
        d.bounds.allocate(6);
        d.keys.allocate(T);
        d.keys_sorted.allocate(T);
        d.ids.allocate(T);
        d.ids_sorted.allocate(T);
        d.bvh.allocate(std::max<size_t>(1, T - 1));
        d.parent_internal.allocate(std::max<size_t>(1, T - 1));
        d.parent_leaf.allocate(T);
        d.counters.allocate(std::max<size_t>(1, T - 1));
        d.refit_ready.allocate(std::max<size_t>(1, T - 1));
        d.refit_ready_count.allocate(1);
        size_t temp = 0;
        HIP_TRY(rocprim::radix_sort_pairs(nullptr, temp, d.keys.get(), d.keys_sorted.get(), d.ids.get(),
                                          d.ids_sorted.get(), T, 0, 30, d.stream));
        d.sort_temp.allocate(std::max<size_t>(temp, 1));
        d.sort_temp_bytes = temp;
        d.cache_count.allocate(d.N);
        HIP_TRY(hipMemset(d.cache_count.get(), 0, d.cache_count.bytes()));
        d.cache_tris.allocate(size_t(kCacheCapacity) * d.N);
        d.cache_box.allocate(6 * size_t(d.N));
        d.cache_pending.allocate(d.N);
        d.cache_pending_count.allocate(2);
        HIP_TRY(hipMemset(d.cache_pending_count.get(), 0, d.cache_pending_count.bytes()));
        d.cache_pending_epoch.allocate(d.N);
        HIP_TRY(hipMemset(d.cache_pending_epoch.get(), 0, d.cache_pending_epoch.bytes()));
        d.cache_snapshot.allocate(1);
        d.swept_epoch = 0;
        d.cache_listed.allocate(d.N);
        d.cache_list.allocate(d.N);
        d.cache_list_count.allocate(1);
        d.best_tri.allocate(d.N);
        d.best_pair.allocate(dev::kPairWords * size_t(d.N));
        rebuild_bvh(false);
    }
    trust_radius_ = std::max(0.25 * model_.maximum_rest_length,
                             8.0 * (max_radius() + material_.collider_offset + material_.barrier_distance));
    HIP_TRY(hipStreamSynchronize(d.stream));
    return KHS_OK;
}

// Morton codes, radix sort, hierarchy and boxes from the current collider.
void Solver::rebuild_bvh(bool swept)
{
    DeviceState &d = *dev_;
    if (!d.T) return;
    d.cache_valid = false;
    const ColliderView cv = d.collider();
    const unsigned tb = grid_for(d.T, kBlock);
    KHS_LAUNCH(d, dev::k_triangle_boxes, dim3(tb), dim3(kBlock), 0, d.stream, cv, swept ? 1 : 0, d.tri_box.get(),
                       d.tri_box_static.get(), d.bound_partials.get());
    KHS_LAUNCH(d, dev::k_bounds_final, dim3(1), dim3(256), 0, d.stream, d.bound_partials.get(), tb, d.bounds.get());
    KHS_LAUNCH(d, dev::k_morton, dim3(tb), dim3(kBlock), 0, d.stream, cv, d.bounds.get(), d.keys.get(),
                       d.ids.get());
    size_t temp = d.sort_temp_bytes;
    HIP_TRY(rocprim::radix_sort_pairs(d.sort_temp.get(), temp, d.keys.get(), d.keys_sorted.get(), d.ids.get(),
                                      d.ids_sorted.get(), d.T, 0, 30, d.stream));
    KHS_LAUNCH(d, dev::k_hierarchy, dim3(tb), dim3(kBlock), 0, d.stream, d.T, d.keys_sorted.get(),
                       d.ids_sorted.get(), d.bvh.get(), d.parent_internal.get(), d.parent_leaf.get(), d.counters.get());
    refit_boxes();
}

// Bottom-up boxes on the current hierarchy (counters zeroed by the caller).
void Solver::refit_boxes()
{
    DeviceState &d = *dev_;
    if (d.T == 1) {
        KHS_LAUNCH(d, dev::k_refit, dim3(1), dim3(kBlock), 0, d.stream, d.T, d.ids_sorted.get(), d.tri_box.get(),
                           d.parent_internal.get(), d.parent_leaf.get(), d.bvh.get(), d.counters.get());
    } else {
        KHS_LAUNCH(d, dev::k_fill_u32, dim3(1), dim3(64), 0, d.stream, d.refit_ready_count.get(), size_t(1), 0u);
        KHS_LAUNCH(d, dev::k_refit_local, dim3(grid_for(d.T, dev::kRefitChunk)), dim3(dev::kRefitChunk), 0,
                           d.stream, d.T, d.ids_sorted.get(), d.tri_box.get(), d.parent_internal.get(),
                           d.parent_leaf.get(), d.bvh.get(), d.counters.get(), d.refit_ready.get(),
                           d.refit_ready_count.get());
        KHS_LAUNCH(d, dev::k_refit_top, dim3(1), dim3(1024), 0, d.stream, d.parent_internal.get(), d.bvh.get(),
                           d.counters.get(), d.refit_ready.get(), d.refit_ready_count.get());
    }
    HIP_TRY(hipGetLastError());
}

// New boxes (static or swept) on the current hierarchy.
void Solver::refit_bvh(bool swept)
{
    DeviceState &d = *dev_;
    if (!d.T) return;
    d.cache_valid = false;
    const unsigned tb = grid_for(d.T, kBlock);
    KHS_LAUNCH(d, dev::k_triangle_boxes, dim3(tb), dim3(kBlock), 0, d.stream, d.collider(), swept ? 1 : 0,
                       d.tri_box.get(), d.tri_box_static.get(), d.bound_partials.get());
    KHS_LAUNCH(d, dev::k_zero_counters, dim3(tb), dim3(kBlock), 0, d.stream, d.counters.get(),
                       d.T > 1 ? d.T - 1 : 1u);
    refit_boxes();
}

// Brings the candidate lists up to date for a query pass (all of them
// after the collider boxes changed): elements whose query box left their
// cached box gather again.
void Solver::refresh_contact_cache(int mode, bool full, bool swept_leaves, bool with_roots)
{
    DeviceState &d = *dev_;
    // the largest pad of any element's queries (per-strand radii included)
    const double contact_pad = max_radius() + material_.collider_offset + material_.barrier_distance;
    const double sweep_target = max_radius() + material_.collider_offset + desc_.minimum_gap;
    const double pad = std::fmax(contact_pad, sweep_target);
    // Half the pad: measured against 0.35-2 pads on the groom, a thinner
    // skin shortens every list pass by more than the extra regathers cost.
    d.cache_skin = 0.5 * pad;
    const ContactCache cc = d.cache();
    if (full) {
        d.swept_covers = false;
        KHS_LAUNCH(d, dev::k_fill_u32, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.cache_listed.get(),
                           size_t(d.N), 0u);
        KHS_LAUNCH(d, dev::k_fill_u32, dim3(1), dim3(64), 0, d.stream, d.cache_list_count.get(), size_t(1), 0u);
    }
    const dim3 grid(d.node_blocks()), block(kBlock);
    if (full) {
        // every element gathers: one lane each
        KHS_LAUNCH(d, dev::k_contact_refresh<dev::kAtState>, grid, block, 0, d.stream, d.rods(),
                           swept_leaves ? d.swept_collider() : d.collider(), cc, d.q.get(), d.q.get(),
                           d.dir.get(), d.alpha.get(), pad, 1);
    } else {
        // the few elements that left their box gather, each with its
        // wavefront
        const ColliderView col = d.collider();
        if (mode == dev::kAtState)
            KHS_LAUNCH(d, dev::k_contact_check<dev::kAtState>, grid, block, 0, d.stream, d.rods(), col, cc,
                               d.q.get(), d.q.get(), d.dir.get(), d.alpha.get(), dev::TrialStep{}, pad,
                               with_roots ? 1 : 0, (const double *)nullptr, (Partial *)nullptr);
        else if (mode == dev::kAtTrial) {
            KHS_LAUNCH(d, dev::k_contact_check<dev::kAtTrial>, grid, block, 0, d.stream, d.rods(), col, cc,
                               d.q.get(), d.q.get(), d.dir.get(), d.alpha.get(), d.trial_step, pad,
                               with_roots ? 1 : 0, (const double *)nullptr, (Partial *)nullptr);
            d.trial_step.source = dev::kAlphaStored;
        }
        else {
            // The swept pass queues its elements, marked with a new epoch,
            // for the gathers of the step-bound kernel after it
            // (swept_bound_async).
            if (++d.swept_epoch == 0) {
                KHS_LAUNCH(d, dev::k_fill_u32, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream,
                                   d.cache_pending_epoch.get(), size_t(d.N), 0u);
                d.swept_epoch = 1;
            }
            KHS_LAUNCH(d, dev::k_contact_check<dev::kSwept>, grid, block, 0, d.stream, d.rods(), col, d.cache(),
                               d.q.get(), d.q.get(), d.dir.get(), d.alpha.get(), dev::TrialStep{}, pad,
                               with_roots ? 1 : 0, (const double *)d.grad.get(), d.partials.get());
        }
    }
    HIP_TRY(hipGetLastError());
    d.cache_valid = true;
}

// The line-search step bound along dir over the listed elements: the swept
// check, then one kernel that gathers the elements it queued (bounding each
// where it gathers it) and bounds the others from their lists; the two
// pending counters alternate, each cleared by the kernel before its next
// use. Leaves kSweptBoundBlocks per-block steps in list_alpha.
void Solver::swept_bound_async(const StepParams &cp, const unsigned *stop)
{
    DeviceState &d = *dev_;
    refresh_contact_cache(dev::kSwept, false);
    d.swept_covers = true;
    KHS_LAUNCH(d, dev::k_swept_ccd, dim3(kSweptBoundBlocks), dim3(kBlock), 0, d.stream, d.flags.get(),
                       d.collider(), d.cache(), cp, d.q.get(), d.dir.get(), d.list_alpha.get(), d.ccd_bound.get(),
                       int(d.ccd_slot), d.cache_pending_count.get() + (d.pending_slot ^ 1u), stop);
    HIP_TRY(hipGetLastError());
    d.ccd_slot ^= 1u;
    d.pending_slot ^= 1u;
}

// Closest triangle of every listed element at q (or at alpha d + q_base
// with the device's alpha), one wavefront per element; partials (may be
// null) receive the candidate count. Elements with a fixed left node are
// refreshed and included only for the friction pass (`with_roots`).
void Solver::closest_pass(bool line_search, unsigned long long *block_candidates, bool with_roots)
{
    DeviceState &d = *dev_;
    const StepParams sp = step_params(1.0);
    // A trial's query boxes lie inside the boxes the swept check made every
    // list cover (each trial point alpha d + q_base, 0 < alpha <= 1, lies
    // between q_base and q_base + d in every coordinate, rounding included),
    // and the coverage test is monotone: its check would gather nothing.
    const bool covered = contact_reuse_ && line_search && d.swept_covers && !with_roots;
    if (!covered) refresh_contact_cache(line_search ? dev::kAtTrial : dev::kAtState, false, false, with_roots);
    const ContactCache cc = d.cache();
    if (line_search) {
        KHS_LAUNCH(d, dev::k_list_closest<dev::kAtTrial>, dim3(kPairBlocks), dim3(kBlock), 0, d.stream,
                           d.flags.get(), d.collider(), cc, sp, d.q.get(), d.q.get(), d.dir.get(),
                           d.alpha.get(), d.trial_step, d.best_tri.get(), d.best_pair.get(), block_candidates,
                           with_roots ? d.root_elements.get() : nullptr, with_roots ? d.root_element_count : 0u,
                           d.skip());
        d.trial_step.source = dev::kAlphaStored;
    } else {
        KHS_LAUNCH(d, dev::k_list_closest<dev::kAtState>, dim3(kPairBlocks), dim3(kBlock), 0, d.stream,
                           d.flags.get(), d.collider(), cc, sp, d.q.get(), d.q.get(), d.dir.get(),
                           d.alpha.get(), dev::TrialStep{}, d.best_tri.get(), d.best_pair.get(), block_candidates,
                           with_roots ? d.root_elements.get() : nullptr, with_roots ? d.root_element_count : 0u,
                           (const unsigned *)nullptr);
    }
    HIP_TRY(hipGetLastError());
}

KhsResult Solver::initial_feasibility()
{
    // reference solver.cpp:634-657
    DeviceState &d = *dev_;
    Buffer<double> dpos;
    dpos.allocate(3 * size_t(d.N));
    KHS_LAUNCH(d, dev::k_gather_positions, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.q.get(),
                       (const uint32_t *)nullptr, d.N, dpos.get());
    HIP_TRY(hipGetLastError());
    const RodView rv = d.rods();
    const ColliderView cv = d.collider();
    if (collider_.closed) {
        double x_max = -kInf;
        for (const Vec3 &v : collider_.vertices) x_max = std::max(x_max, v.x);
        const unsigned init = 0xffffffffu;
        HIP_TRY(hipMemcpyAsync(d.flag.get(), &init, sizeof(unsigned), hipMemcpyHostToDevice, d.stream));
        KHS_LAUNCH(d, dev::k_point_inside, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, rv, cv, dpos.get(),
                           x_max, d.flag.get());
        HIP_TRY(hipGetLastError());
        unsigned inside = 0;
        HIP_TRY(hipMemcpyAsync(&inside, d.flag.get(), sizeof(unsigned), hipMemcpyDeviceToHost, d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
        if (inside != 0xffffffffu) {
            error_ = "Initial intersection: internal node " + std::to_string(inside) + " is inside the closed collider.";
            return KHS_ERROR_INITIAL_INTERSECTION;
        }
    }
    const StepParams p = step_params(1.0);
    KHS_LAUNCH(d, dev::k_initial_gap, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, rv, cv, p, dpos.get(),
                       d.partials.get());
    HIP_TRY(hipGetLastError());
    const Partial r = d.reduce(d.node_blocks());
    build_stats_.initial_minimum_gap = r.min_gap;
    if (r.min_gap <= desc_.minimum_gap) {
        std::ostringstream ss;
        ss << "Initial intersection or negative gap: minimum gap " << r.min_gap << " m, " << r.candidates
           << " candidates.";
        error_ = ss.str();
        return KHS_ERROR_INITIAL_INTERSECTION;
    }
    return KHS_OK;
}

KhsResult Solver::update_runtime_parameters(const KhsSolverDesc &desc, const KhsHairMaterial &material)
{
    // reference solver.cpp:558-606, cuda_backend.cu:409-430
    if (!built_) {
        error_ = "Build the solver before updating runtime parameters.";
        return KHS_ERROR_INVALID_STATE;
    }
    if (desc.maximum_element_length != desc_.maximum_element_length ||
        desc.minimum_dynamic_length != desc_.minimum_dynamic_length || desc.fixed_root_nodes != desc_.fixed_root_nodes) {
        error_ = "maximum_element_length, minimum_dynamic_length and fixed_root_nodes cannot change after the build.";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    desc_ = desc;
    material_ = material;
    apply_strand_materials();  // masses and trust radius of the new material
    error_.clear();
    return KHS_OK;
}

KhsResult Solver::update_collider(const KhsVec3 *vertices, uint32_t vertex_count)
{
    // reference solver.cpp:683-708
    if (!built_ || !has_collider_) {
        error_ = "No built collider.";
        return KHS_ERROR_INVALID_STATE;
    }
    if (!vertices || vertex_count != collider_.vertices.size()) {
        error_ = "Collider vertex count changed.";
        return KHS_ERROR_INVALID_COLLIDER;
    }
    std::vector<Vec3> v = to_vec3(vertices, vertex_count);
    for (const Vec3 &x : v)
        if (!is_finite(x)) {
            error_ = "Collider vertex is not finite.";
            return KHS_ERROR_INVALID_COLLIDER;
        }
    const std::vector<double> flat = flatten(v);
    dev_->col_pend.upload(flat.data(), flat.size());
    collider_pending_host_ = std::move(v);
    return KHS_OK;
}

KhsResult Solver::update_roots(const KhsVec3 *points, uint32_t point_count)
{
    // reference solver.cpp:710-742
    if (!built_ || !points || point_count != points_.size()) {
        error_ = "Root target count does not match the hair curves.";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    std::vector<Vec3> p = to_vec3(points, point_count);
    for (const Vec3 &x : p)
        if (!is_finite(x)) {
            error_ = "Root target is not finite.";
            return KHS_ERROR_INVALID_HAIR;
        }
    if (!device_root_targets(p, dev_->root_pend.get(), dev_->rot_pend.get())) {
        error_ = "Cannot build root orientation from the targets.";
        return KHS_ERROR_INVALID_HAIR;
    }
    return KHS_OK;
}

// Root targets of one frame of original points on the device, into pos and
// rot (3N each, device); false when a tangent is degenerate.
bool Solver::device_root_targets(const std::vector<Vec3> &points, double *pos, double *rot)
{
    DeviceState &d = *dev_;
    const std::vector<double> flat = flatten(points);
    d.root_points.upload(flat.data(), flat.size());
    KHS_LAUNCH(d, dev::k_fill_u32, dim3(1), dim3(64), 0, d.stream, d.root_flag.get(), size_t(1), 0u);
    KHS_LAUNCH(d, dev::k_root_positions, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.N,
                       (const uint32_t *)d.bind_left.get(), (const uint32_t *)d.bind_right.get(),
                       (const double *)d.bind_t.get(), (const double *)d.root_points.get(), pos);
    KHS_LAUNCH(d, dev::k_root_rotations, dim3(grid_for(d.S, kBlock)), dim3(kBlock), 0, d.stream, d.rods(),
                       (const double *)pos, rot, d.root_flag.get());
    HIP_TRY(hipGetLastError());
    unsigned bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, d.root_flag.get(), sizeof(bad), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    return bad == 0;
}

// -------------------------------------------------------------- animation
KhsResult Solver::allocate_animation(uint32_t frames)
{
    if (!built_) return KHS_ERROR_INVALID_STATE;
    if (frames < 1) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid animation frame count.");
    try {
        DeviceState &d = *dev_;
        d.root_anim.allocate(size_t(frames) * 3 * d.N);
        d.rot_anim.allocate(size_t(frames) * 3 * d.N);
        if (d.V) d.col_anim.allocate(size_t(frames) * 3 * d.V);
    } catch (const std::exception &e) {
        return fail(KHS_ERROR_OUT_OF_MEMORY, e.what());
    }
    animation_frames_ = frames;
    std::fill(ring_frame_.begin(), ring_frame_.end(), ~0u);
    root_frame_set_.assign(frames, false);
    collider_frame_set_.assign(frames, !has_collider_);
    gpu_stats_.animation_frame_count = frames;
    gpu_stats_.resident_bytes = dev_->resident_bytes();
    return KHS_OK;
}

KhsResult Solver::set_root_animation_frame(uint32_t frame, const KhsVec3 *points, uint32_t point_count)
{
    if (!built_ || !points || point_count != points_.size()) return KHS_ERROR_INVALID_ARGUMENT;
    std::vector<Vec3> p = to_vec3(points, point_count);
    for (const Vec3 &x : p)
        if (!is_finite(x)) return KHS_ERROR_INVALID_HAIR;
    if (frame >= animation_frames_) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid root animation frame.");
    const size_t n3 = 3 * size_t(dev_->N);
    if (!device_root_targets(p, dev_->root_anim.get() + size_t(frame) * n3, dev_->rot_anim.get() + size_t(frame) * n3))
        return KHS_ERROR_INVALID_HAIR;
    root_frame_set_[frame] = true;
    return KHS_OK;
}

KhsResult Solver::set_collider_animation_frame(uint32_t frame, const KhsVec3 *vertices, uint32_t vertex_count)
{
    if (!built_) return KHS_ERROR_INVALID_STATE;
    const uint32_t vc = dev_->V;
    if (frame >= animation_frames_ || vertex_count != vc || (vertex_count && !vertices))
        return fail(KHS_ERROR_INVALID_COLLIDER, "Invalid collider animation frame.");
    if (vc) {
        const std::vector<double> flat = flatten(to_vec3(vertices, vertex_count));
        dev_->col_anim.upload(flat.data(), flat.size(), size_t(frame) * 3 * vc);
    }
    collider_frame_set_[frame] = true;
    return KHS_OK;
}

KhsResult Solver::finalize_animation()
{
    if (!built_) return KHS_ERROR_INVALID_STATE;
    for (bool b : root_frame_set_)
        if (!b) return fail(KHS_ERROR_INVALID_STATE, "Animation has missing frames.");
    for (bool b : collider_frame_set_)
        if (!b) return fail(KHS_ERROR_INVALID_STATE, "Animation has missing frames.");
    if (animation_frames_ < 1) return fail(KHS_ERROR_INVALID_STATE, "Animation has missing frames.");
    animation_ready_ = true;
    current_frame_ = 0;
    return KHS_OK;
}

// ---------------------------------------------------------------- solving
// Enqueues one evaluation of the incremental potential (the line-search
// variant at alpha d + q_base, alpha read from the device) and its
// reduction into reduced[kGradSlot] (with the gradient) or reduced[kEvalSlot].
// The scalars of a substep of length dt (material weights as in the
// reference's element kernel, cuda_backend.cu:905-910).
StepParams Solver::step_params(double dt) const
{
    StepParams p{};
    p.dt = dt;
    p.inv_dt = 1.0 / dt;
    p.inv_dt2 = p.inv_dt * p.inv_dt;
    p.gravity[0] = desc_.gravity.x;
    p.gravity[1] = desc_.gravity.y;
    p.gravity[2] = desc_.gravity.z;
    p.radius_offset = material_.radius + material_.collider_offset;
    p.barrier = material_.barrier_distance;
    p.stiffness = material_.contact_stiffness;
    p.damping = material_.mass_damping;
    p.friction = material_.friction;
    p.smoothing = material_.friction_smoothing;
    p.minimum_gap = desc_.minimum_gap;
    const double r = material_.radius;
    const double area = kPi * r * r, inertia = kPi * r * r * r * r / 4.0;
    const double G = material_.young_modulus / (2.0 * (1.0 + material_.poisson_ratio));
    p.weights[0] = material_.shear_correction * G * area;
    p.weights[1] = material_.shear_correction * G * area;
    p.weights[2] = material_.young_modulus * area;
    p.weights[3] = material_.young_modulus * inertia;
    p.weights[4] = material_.young_modulus * inertia;
    p.weights[5] = 2.0 * G * inertia;
    p.contact_pad = material_.radius + material_.collider_offset + material_.barrier_distance;
    p.sweep_target = material_.radius + material_.collider_offset + desc_.minimum_gap;
    p.has_collider = dev_ && dev_->T > 0;
    p.material = (dev_ && dev_->elem_material.size()) ? dev_->elem_material.get() : nullptr;
    p.collider_offset = material_.collider_offset;
    p.poisson = material_.poisson_ratio;
    p.shear_correction = material_.shear_correction;
    return p;
}

void Solver::evaluate_async(bool gradient, double dt, bool line_search, unsigned *stop, double abs_tol,
                            double rel_tol, double initial)
{
    DeviceState &d = *dev_;
    const StepParams p = step_params(dt);
    const unsigned blocks = d.eval_blocks();
    const RodView rv = d.rods();
    const ColliderView cv = d.collider();
    const ContactCache cc = d.cache();
    hipEvent_t t0 = d.tic();
    // With current candidate lists the list pass finds the closest
    // triangles first; its partials follow the evaluation's. At an
    // accepted trial's iterate its own list pass already did.
    const bool list_pass = d.cache_valid;
    if (list_pass && !(gradient && d.closest_at_q)) closest_pass(line_search, d.list_words.get());
    d.closest_at_q = false;
    const uint32_t *best = list_pass ? d.best_tri.get() : nullptr;
    // A line-search trial reads its base q and writes alpha d + q into
    // q_trial; the host swaps the two when it accepts the trial.
    const bool trial_out = line_search && !gradient;
    // the variants without the tree walk when the list pass answered every query
    auto launch = [&](auto kernel, const unsigned *stop) {
        KHS_LAUNCH(d, kernel, dim3(blocks), dim3(kBlock), 0, d.stream, rv, cv, cc, best,
                           (const double *)d.best_pair.get(), p, trial_out ? d.q_trial.get() : d.q.get(),
                           d.q_old.get(), d.vel.get(), d.q.get(), d.dir.get(), d.alpha.get(), d.grad.get(),
                           d.Dblk.get(), d.Ublk.get(), d.partials.get(), stop);
    };
    // variants with per-strand stiffness weights only when they are set
    const bool mat = d.elem_material.size() != 0;
    if (gradient) {
        if (list_pass) launch(mat ? dev::k_evaluate<true, false, true, true> : dev::k_evaluate<true, false, true, false>,
                              (const unsigned *)nullptr);
        else launch(mat ? dev::k_evaluate<true, false, false, true> : dev::k_evaluate<true, false, false, false>,
                    (const unsigned *)nullptr);
    } else if (line_search) {
        if (list_pass) launch(mat ? dev::k_evaluate<false, true, true, true> : dev::k_evaluate<false, true, true, false>,
                              d.skip());
        else launch(mat ? dev::k_evaluate<false, true, false, true> : dev::k_evaluate<false, true, false, false>,
                    d.skip());
    } else {
        launch(mat ? dev::k_evaluate<false, false, false, true> : dev::k_evaluate<false, false, false, false>,
               (const unsigned *)nullptr);
    }
    HIP_TRY(hipGetLastError());
    if (gradient) d.toc(t0, d.totals.assembly_ms, d.totals.assembly_launches);
    else d.toc(t0, d.totals.energy_ms, d.totals.energy_launches);
    d.reduce_async(blocks, gradient ? DeviceState::kGradSlot : DeviceState::kEvalSlot,
                   dev::ConvergenceTest{abs_tol, rel_tol, initial, stop},
                   dev::ListPartials{list_pass ? d.list_words.get() : nullptr, nullptr, list_pass ? kPairBlocks : 0u});
}

Solver::Reduced Solver::unpack(const Partial &r2)
{
    Reduced out{};
    out.objective = r2.sum[kObjective];
    out.elastic = r2.sum[kElastic];
    out.contact = r2.sum[kContact];
    out.kinetic = r2.sum[kKinetic];
    out.grad_norm2 = r2.sum[kGradNorm2];
    out.dir_norm2 = r2.sum[kDirNorm2];
    out.min_gap = r2.min_gap;
    out.candidates = r2.candidates;
    out.active = r2.active;
    out.feasible = r2.infeasible == 0;
    return out;
}

Solver::Reduced Solver::evaluate(bool gradient, double dt, bool line_search)
{
    evaluate_async(gradient, dt, line_search);
    const DeviceState::HostView v = dev_->read_back();
    return unpack(gradient ? v.grad : v.eval);
}

// Steepest descent and its trust-radius factor when the choice is not the
// block direction (reference cuda_backend.cu:623-638).
void Solver::steepest_direction_async()
{
    DeviceState &d = *dev_;
    d.swept_covers = false;
    const RodView rv = d.rods();
    const size_t dofs = 6 * size_t(d.N);
    KHS_LAUNCH(d, dev::k_steepest, dim3(grid_for(dofs, kBlock)), dim3(kBlock), 0, d.stream, rv, d.choice(),
                       d.grad.get(), d.precond.get(), d.dir.get());
    KHS_LAUNCH(d, dev::k_direction_stats, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, rv, d.grad.get(),
                       d.dir.get(), (const unsigned *)nullptr, (const DirectionChoice *)d.choice(),
                       d.partials.get());
    d.reduce_async(d.node_blocks(), DeviceState::kSteepSlot);
    KHS_LAUNCH(d, dev::k_choose_steepest_scale, dim3(1), dim3(1), 0, d.stream,
                       (const Partial *)(d.reduced.get() + DeviceState::kSteepSlot), trust_radius_, d.choice(),
                       d.mirror_choice());
    KHS_LAUNCH(d, dev::k_scale_direction, dim3(grid_for(dofs, kBlock)), dim3(kBlock), 0, d.stream,
                       (const DirectionChoice *)d.choice(), 0, d.dir.get(), dofs);
    HIP_TRY(hipGetLastError());
}

// Sets the next trial's step: the CCD bound of the last reduction or a
// host value. With contact lists the trial's first kernel (the list check)
// resolves it; otherwise a one-thread kernel does.
void Solver::set_trial_step(int source, double value)
{
    DeviceState &d = *dev_;
    const Partial *ccd = (const Partial *)(d.reduced.get() + DeviceState::kCcdSlot);
    if (d.cache_valid) {
        d.trial_step = dev::TrialStep{ccd, value, source, d.T ? 1 : 0};
        return;
    }
    if (source == dev::kAlphaFromCcd)
        KHS_LAUNCH(d, dev::k_choose_alpha, dim3(1), dim3(1), 0, d.stream, ccd, d.T ? 1 : 0, d.alpha.get());
    else
        KHS_LAUNCH(d, dev::k_set_double, dim3(1), dim3(1), 0, d.stream, d.alpha.get(), value);
    HIP_TRY(hipGetLastError());
}

// One block-tridiagonal solve on the stream: block Thomas (FP64, or FP32
// in precision mode 1) or, in FP64, the twisted factorization (solve
// algorithm 1) or the fused elimination (2; block Thomas where D, U and g
// together reach kMaxBufferBytes, about 6.8M nodes). Returns the grid
// size, the count of Partials it leaves.
uint32_t Solver::launch_block_solve(unsigned *failed, unsigned *next_failed, const unsigned *stop)
{
    DeviceState &d = *dev_;
    const RodView rv = d.rods();
    uint32_t blocks = d.solve_blocks();
    if (precision_mode_ == 1) {
        KHS_LAUNCH(d, dev::k_block_solve<float>, dim3(blocks), dim3(kWave), kSolveLdsReserve, d.stream, rv,
                           d.solve_order.get(), d.Dblk.get(), d.Ublk.get(), d.grad.get(), d.Cp.get(), d.dp.get(),
                           d.dir.get(), failed, next_failed, d.partials.get(), stop);
    } else if (solve_algorithm_ == 1) {
        blocks = d.twisted_blocks();
        KHS_LAUNCH(d, dev::k_block_solve_twisted, dim3(blocks), dim3(kWave), kSolveLdsReserve, d.stream, rv,
                           d.solve_order.get(), d.Dblk.get(), d.Ublk.get(), d.grad.get(), d.Cp.get(), d.dp.get(),
                           d.dir.get(), failed, next_failed, d.partials.get(), stop);
    } else if (solve_algorithm_ == 2 && d.solve_in.bytes() < dev::kMaxBufferBytes) {
        // its factor window is the LDS reservation (static, 38.9 KB)
        KHS_LAUNCH(d, dev::k_block_solve_fused, dim3(blocks), dim3(kWave), 0, d.stream, rv,
                           d.solve_order.get(), (const double *)d.solve_in.get(), d.Cp.get(), d.dp.get(), d.dir.get(),
                           failed, next_failed, d.partials.get(), stop);
    } else {
        KHS_LAUNCH(d, dev::k_block_solve<double>, dim3(blocks), dim3(kWave), kSolveLdsReserve, d.stream, rv,
                           d.solve_order.get(), d.Dblk.get(), d.Ublk.get(), d.grad.get(), d.Cp.get(), d.dp.get(),
                           d.dir.get(), failed, next_failed, d.partials.get(), stop);
    }
    HIP_TRY(hipGetLastError());
    return blocks;
}

// Block direction limited to the trust radius, with the choice whether it
// is usable (solved and descending) made on the device (reference
// cuda_backend.cu:623-648). The choice reaches the host with the next
// read-back; when the direction is not usable the host replaces it by
// steepest descent and redoes the step bound (optimize_substep).
void Solver::make_direction_async()
{
    DeviceState &d = *dev_;
    d.swept_covers = false;
    const RodView rv = d.rods();
    unsigned *failed = d.solve_failed.get() + d.failed_slot;
    hipEvent_t t0 = d.tic();
    const uint32_t blocks = launch_block_solve(failed, d.solve_failed.get() + (d.failed_slot ^ 1u), d.skip());
    d.failed_slot ^= 1u;
    d.toc(t0, d.totals.solve_ms, d.totals.solve_launches);
    // the solve leaves the direction's statistics in one partial per block
    d.reduce_async(blocks, DeviceState::kDirSlot);
    // the choice and the trust scaling (the line search's base is q itself)
    const size_t dofs = 6 * size_t(d.N);
    KHS_LAUNCH(d, dev::k_apply_direction_choice, dim3(grid_for(dofs, kBlock)), dim3(kBlock), 0, d.stream,
                       (const Partial *)(d.reduced.get() + DeviceState::kDirSlot), trust_radius_, d.choice(),
                       d.mirror_choice(), d.dir.get(), dofs);
    HIP_TRY(hipGetLastError());
}

void Solver::direction_async()
{
    if (direction_mode_ == 0) {
        make_direction_async();
        return;
    }
    DeviceState &d = *dev_;
    d.swept_covers = false;
    KHS_LAUNCH(d, dev::k_force_steepest, dim3(1), dim3(1), 0, d.stream, d.choice(), d.mirror_choice());
    HIP_TRY(hipGetLastError());
    steepest_direction_async();
}

// Jacobi diagonal of the substep (reference cuda_backend.cu:815-841).
void Solver::launch_preconditioner(double dt)
{
    DeviceState &d = *dev_;
    const double r = material_.radius;
    const double area = kPi * r * r, inertia = kPi * r * r * r * r / 4.0;
    const double G = material_.young_modulus / (2.0 * (1.0 + material_.poisson_ratio));
    const double tk = std::fmax(material_.young_modulus * area, material_.shear_correction * G * area);
    const double ra = material_.shear_correction * G * area;
    const double rb = std::fmax(material_.young_modulus * inertia, 2.0 * G * inertia);
    KHS_LAUNCH(d, dev::k_preconditioner, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.rods(), 1.0 / dt,
                       material_.mass_damping, tk, ra, rb,
                       d.elem_material.size() ? (const double *)d.elem_material.get() : nullptr,
                       material_.poisson_ratio, material_.shear_correction, d.precond.get());
    HIP_TRY(hipGetLastError());
}

KhsResult Solver::optimize_substep(double dt)
{
    // reference cuda_backend.cu:650-688
    DeviceState &d = *dev_;
    const RodView rv = d.rods();
    launch_preconditioner(dt);
    const double dofs = double(6 * size_t(d.N));
    const double abs_tol = convergence_mode_ ? desc_.absolute_tolerance
                                             : std::fmax(desc_.absolute_tolerance, 1.0e-5 * std::sqrt(dofs));
    const double rel_tol = convergence_mode_ ? desc_.relative_tolerance : std::fmax(desc_.relative_tolerance, 1.0e-4);
    // Work enqueued ahead for the next Newton iteration skips itself on the
    // device when the gradient before it has converged.
    d.skipping = true;
    struct SkipOff {
        DeviceState &d;
        ~SkipOff() { d.skipping = false; }
    } skip_off{d};
    evaluate_async(true, dt, false, d.stop.get(), abs_tol, rel_tol, -1.0);
    Reduced current = unpack(d.read_back().grad);
    if (!current.feasible) return fail(KHS_ERROR_NOT_CONVERGED, "State is outside the barrier feasible region.");
    const double initial = std::sqrt(current.grad_norm2);
    step_stats_.initial_residual_norm = std::fmax(step_stats_.initial_residual_norm, initial);
    const StepParams cp = step_params(dt);
    // The step bound along the current direction and the first trial at it.
    auto bound_and_trial = [&]() {
        phase_.store(KHS_PHASE_CCD);
        hipEvent_t t0 = d.tic();
        const bool list_pass = d.cache_valid;
        // With lists the swept check leaves the slope's partials and the
        // bound kernel bounds every listed element (overflowed lists too);
        // without, one pass takes the slope and walks the tree per element.
        if (list_pass) swept_bound_async(cp, d.skip());
        else
            KHS_LAUNCH(d, dev::k_ccd_slope, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, rv, d.collider(), cp,
                               d.q.get(), d.grad.get(), d.dir.get(), d.partials.get());
        HIP_TRY(hipGetLastError());
        d.toc(t0, d.totals.collision_ms, d.totals.collision_launches);
        d.reduce_async(d.node_blocks(), DeviceState::kCcdSlot, dev::ConvergenceTest{0.0, 0.0, 0.0, nullptr},
                       dev::ListPartials{nullptr, list_pass ? d.list_alpha.get() : nullptr,
                                         list_pass ? kSweptBoundBlocks : 0u});
        set_trial_step(dev::kAlphaFromCcd, 0.0);
        // The first trial runs at the device's step before the host has
        // seen the bound; one read-back returns direction choice, bound and
        // trial. A trial the host would not have run is discarded, and every
        // path that does not accept leaves q as it is (trials write q_trial).
        phase_.store(KHS_PHASE_LINE_SEARCH);
        evaluate_async(false, dt, true);
    };
    bool head_ready = false;  // direction, bound and first trial already enqueued
    DeviceState::HostView ahead{};
    for (uint32_t it = 0; it <= desc_.newton_iterations; ++it) {
        iteration_.store(it);
        if (cancel_.load()) return fail(KHS_ERROR_INVALID_STATE, "Solve cancelled.");
        const double residual = std::sqrt(current.grad_norm2);
        const double relative = residual / std::fmax(initial, 1e-30);
        step_stats_.final_residual_norm = residual;
        step_stats_.relative_residual_norm = relative;
        step_stats_.minimum_gap = std::fmin(step_stats_.minimum_gap, current.min_gap);
        step_stats_.contact_candidate_count += current.candidates;
        step_stats_.active_contact_count += current.active;
        step_stats_.kinetic_energy = current.kinetic;
        step_stats_.elastic_energy = current.elastic;
        step_stats_.contact_energy = current.contact;
        step_stats_.friction_energy = 0.0;
        if (residual <= abs_tol || relative <= rel_tol) {
            // work enqueued ahead skipped itself (the device made the same
            // test), so q is still this iterate
            KHS_LAUNCH(d, dev::k_update_velocity, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, rv,
                               d.q.get(), d.q_old.get(), d.vel.get(), 1.0 / dt);
            HIP_TRY(hipGetLastError());
            return KHS_OK;
        }
        if (it == desc_.newton_iterations) break;
        phase_.store(KHS_PHASE_LINEAR_SOLVE);
        DeviceState::HostView v{};
        bool have_view = false;
        if (head_ready) {
            v = ahead;
            have_view = true;
            head_ready = false;
        } else {
            direction_async();
        }
        ++step_stats_.linear_solves;
        // with the forced rule the first direction already is steepest descent
        bool accepted = false, block_direction = direction_mode_ == 0, counted = false;
        double step_dir_norm2 = 0.0;  // |d|^2 of the accepted trial's direction
        double alpha = 1.0;
        for (int attempt = 0; attempt < 2;) {
            if (!have_view) {
                bound_and_trial();
                v = d.read_back();
            }
            have_view = false;
            if (attempt == 0 && block_direction && !v.choice.used) {
                // The block direction failed or does not descend: the step
                // starts from steepest descent instead, with no retry.
                steepest_direction_async();
                block_direction = false;
                continue;
            }
            if (!counted) {
                ++(block_direction ? block_directions_ : steepest_directions_);
                counted = true;
            }
            const Partial &cr = v.ccd;
            const double slope = cr.sum[kSlope];
            alpha = 1.0;
            if (d.T) {
                alpha = cr.alpha;
                if (alpha < 1.0) alpha *= 0.9;
                step_stats_.ccd_step_limit = std::fmin(step_stats_.ccd_step_limit, alpha);
            }
            accepted = false;
            if (alpha >= desc_.minimum_line_search_step) {
                Reduced trial = unpack(v.eval);
                for (uint32_t search = 0; search < desc_.line_search_iterations; ++search) {
                    if (search > 0) {
                        set_trial_step(dev::kAlphaGiven, alpha);
                        trial = evaluate(false, dt, true);
                    }
                    ++step_stats_.line_search_evaluations;
                    if (trial.feasible && std::isfinite(trial.objective) &&
                        trial.objective <= current.objective + 1e-4 * alpha * slope) {
                        accepted = true;
                        step_dir_norm2 = trial.dir_norm2;
                        d.q.swap(d.q_trial);  // the trial is the new iterate
                        break;
                    }
                    alpha *= 0.5;
                    if (alpha < desc_.minimum_line_search_step) break;
                }
            }
            if (accepted) break;
            if (!block_direction) break;
            KHS_LAUNCH(d, dev::k_force_steepest, dim3(1), dim3(1), 0, d.stream, d.choice(), d.mirror_choice());
            steepest_direction_async();
            ++steepest_directions_;
            block_direction = false;
            ++attempt;
        }
        if (!accepted) {
            return fail(KHS_ERROR_NOT_CONVERGED, "Gauss-Newton line search failed.");
        }
        // The gradient at the accepted iterate (whose closest triangles the
        // trial's list pass left); unless the budget ends here, the next
        // iteration's direction, bound and first trial follow before the one
        // wait (they skip themselves if this gradient converged).
        d.closest_at_q = contact_reuse_;
        evaluate_async(true, dt, false, d.stop.get(), abs_tol, rel_tol, initial);
        if (it + 1 < desc_.newton_iterations) {
            direction_async();
            bound_and_trial();
            head_ready = true;
        }
        ahead = d.read_back();
        current = unpack(ahead.grad);
        const double dnorm = std::sqrt(step_dir_norm2);
        step_stats_.increment_norm = alpha * dnorm;
        step_stats_.accepted_step_length = alpha;
        ++step_stats_.newton_iterations;
        phase_.store(KHS_PHASE_ASSEMBLING);
        if (step_stats_.increment_norm / std::sqrt(dofs) <= 2.0 * desc_.increment_tolerance) {
            // (a trial enqueued ahead wrote only q_trial)
            KHS_LAUNCH(d, dev::k_update_velocity, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, rv,
                               d.q.get(), d.q_old.get(), d.vel.get(), 1.0 / dt);
            HIP_TRY(hipGetLastError());
            return KHS_OK;
        }
    }
    KHS_LAUNCH(d, dev::k_update_velocity, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, rv, d.q.get(),
                       d.q_old.get(), d.vel.get(), 1.0 / dt);
    HIP_TRY(hipGetLastError());
    return KHS_OK;
}

void Solver::enqueue_sweep(const StepParams &sp)
{
    DeviceState &d = *dev_;
    rebuild_bvh(true);
    // Every list is gathered against the swept triangle boxes: the sweep
    // below queries those, and since they contain the current boxes the
    // lists also serve the substep's other queries, which filter with the
    // current boxes.
    refresh_contact_cache(dev::kAtState, true, true);
    const unsigned long long none = ~0ull;
    KHS_LAUNCH(d, dev::k_fill_u64, dim3(1), dim3(64), 0, d.stream, d.key64.get(), size_t(1), none);
    KHS_LAUNCH(d, dev::k_list_sweep, dim3(kPairBlocks), dim3(kBlock), 0, d.stream, d.flags.get(),
                       d.strand_of.get(), d.collider(), d.cache(), d.tri_box.get(), sp, d.q.get(), d.key64.get());
    KHS_LAUNCH(d, dev::k_sweep, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.rods(), d.swept_collider(),
                       sp, d.q.get(), (const uint32_t *)d.cache_count.get(), d.key64.get());
    KHS_LAUNCH(d, dev::k_sweep_record, dim3(1), dim3(1), 0, d.stream, d.rods(), d.collider(), sp, d.q.get(),
                       d.key64.get(), d.sweep.get());
    HIP_TRY(hipGetLastError());
}

// Friction impulse of the substep's end state into vel (reference
// cuda_backend.cu:615-621, 941-964).
void Solver::enqueue_friction(const StepParams &sp)
{
    DeviceState &d = *dev_;
    if (d.cache_valid) closest_pass(false, nullptr, true);
    KHS_LAUNCH(d, dev::k_friction, dim3(d.tiled_blocks()), dim3(kBlock), 0, d.stream, d.rods(), d.collider(),
                       d.cache(), d.cache_valid ? d.best_tri.get() : nullptr, sp, d.q.get(), d.vel.get(), d.dv.get());
    KHS_LAUNCH(d, dev::k_apply_dv, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.N, d.dv.get(),
                       d.vel.get());
    HIP_TRY(hipGetLastError());
}

KhsResult Solver::run_frame(double frame_dt, const double *root_a, const double *rot_a, const double *root_b,
                            const double *rot_b, const double *collider_a, const double *collider_b, uint32_t frame,
                            uint32_t frame_count)
{
    // reference cuda_backend.cu:690-721
    if (!(frame_dt > 0 && std::isfinite(frame_dt))) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid time step.");
    DeviceState &d = *dev_;
    const RodView rv = d.rods();
    cancel_.store(false);
    clear_failure();
    frame_clock_.store(std::chrono::steady_clock::now().time_since_epoch().count());
    frame_.store(frame);
    frame_count_.store(frame_count);
    iteration_.store(0);
    HIP_TRY(hipEventRecord(d.frame_begin, d.stream));
    const KhsxKernelTiming before = d.totals;  // phase times of this frame are the differences
    const size_t dof_bytes = d.q.bytes();
    ++d.launch_count;
    HIP_TRY(hipMemcpyAsync(d.snap_q.get(), d.q.get(), dof_bytes, hipMemcpyDeviceToDevice, d.stream));
    ++d.launch_count;
    HIP_TRY(hipMemcpyAsync(d.snap_v.get(), d.vel.get(), dof_bytes, hipMemcpyDeviceToDevice, d.stream));
    uint32_t substeps = desc_.substeps;
    const uint32_t max_sub = d.T ? desc_.maximum_substeps : desc_.substeps;
    uint32_t attempts = 0;
    const size_t col_bytes = d.col_cur.bytes();
    for (;;) {
        ++attempts;
        ++d.launch_count;
        HIP_TRY(hipMemcpyAsync(d.q.get(), d.snap_q.get(), dof_bytes, hipMemcpyDeviceToDevice, d.stream));
        ++d.launch_count;
        HIP_TRY(hipMemcpyAsync(d.vel.get(), d.snap_v.get(), dof_bytes, hipMemcpyDeviceToDevice, d.stream));
        clear_failure();
        error_.clear();
        phase_.store(KHS_PHASE_ASSEMBLING);
        substep_count_.store(substeps);
        iteration_.store(0);
        step_stats_ = {};
        step_stats_.struct_size = sizeof(step_stats_);
        step_stats_.substeps = substeps;
        step_stats_.minimum_gap = kInf;
        step_stats_.ccd_step_limit = 1.0;
        block_directions_ = steepest_directions_ = 0;
        if (d.T) {
            ++d.launch_count;
            HIP_TRY(hipMemcpyAsync(d.col_cur.get(), collider_a, col_bytes, hipMemcpyDeviceToDevice, d.stream));
            if (collider_a != d.col_start.get()) {
                ++d.launch_count;
                HIP_TRY(hipMemcpyAsync(d.col_start.get(), collider_a, col_bytes, hipMemcpyDeviceToDevice, d.stream));
            }
        }
        const double dt = frame_dt / substeps;
        const StepParams sp = step_params(dt);
        bool crossed = false;
        KhsResult result = KHS_OK;
        while (d.substep_events.size() < size_t(substeps) + 1) {
            hipEvent_t e;  // timing only
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            d.substep_events.push_back(e);
        }
        for (uint32_t sub = 1; sub <= substeps; ++sub) {
            substep_.store(sub);
            iteration_.store(0);
            HIP_TRY(hipEventRecord(d.substep_events[sub - 1], d.stream));
            const double f = double(sub) / substeps;
            KHS_LAUNCH(d, dev::k_begin_substep, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, rv, d.q.get(),
                               d.q_old.get(), d.vel.get(), root_a, rot_a, root_b, rot_b, f, dt);
            HIP_TRY(hipGetLastError());
            if (d.T) {
                KHS_LAUNCH(d, dev::k_collider_lerp, dim3(grid_for(3 * size_t(d.V), kBlock)), dim3(kBlock), 0,
                                   d.stream, d.V, d.col_cur.get(), d.col_old.get(), collider_a, collider_b, f);
                hipEvent_t t0 = d.tic();
                phase_.store(KHS_PHASE_CCD);
                enqueue_sweep(sp);
                d.toc(t0, d.totals.collision_ms, d.totals.collision_launches);
                SweepRecord rec{};
                HIP_TRY(hipMemcpyAsync(&rec, d.sweep.get(), sizeof(rec), hipMemcpyDeviceToHost, d.stream));
                d.wait();
                if (rec.claimed) {
                    set_failure_context(KHS_FAILURE_MOVING_COLLIDER_SWEEP, frame, sub, substeps, max_sub, attempts);
                    failure_.strand_index = rec.strand;
                    failure_.element_index = rec.element;
                    failure_.collider_triangle_index = rec.triangle;
                    failure_.distance = rec.distance;
                    failure_.required_distance = rec.required;
                    failure_.clearance = rec.clearance;
                    failure_.collider_substep_displacement = rec.displacement;
                    failure_.collider_frame_displacement = rec.displacement * substeps;
                    failure_.hair_start = {rec.hair_start[0], rec.hair_start[1], rec.hair_start[2]};
                    failure_.hair_end = {rec.hair_end[0], rec.hair_end[1], rec.hair_end[2]};
                    failure_.collider_point = {rec.collider_point[0], rec.collider_point[1], rec.collider_point[2]};
                    crossed = true;
                    break;
                }
                // The swept tree serves the rest of the substep: its boxes
                // contain the current ones, and queries test leaves against
                // the current boxes.
            }
            phase_.store(KHS_PHASE_ASSEMBLING);
            result = optimize_substep(dt);
            if (result != KHS_OK) {
                if (!cancel_.load()) {
                    const uint32_t kind = error_.find("outside the barrier") != std::string::npos
                                              ? KHS_FAILURE_BARRIER_INFEASIBLE
                                              : KHS_FAILURE_NONLINEAR_SOLVE;
                    set_failure_context(kind, frame, sub, substeps, max_sub, attempts);
                }
                break;
            }
            if (d.T && material_.friction > 0.0) enqueue_friction(sp);
            ++step_stats_.converged_substeps;
        }
        if (!crossed && result == KHS_OK) break;
        ++d.launch_count;
        HIP_TRY(hipMemcpyAsync(d.q.get(), d.snap_q.get(), dof_bytes, hipMemcpyDeviceToDevice, d.stream));
        ++d.launch_count;
        HIP_TRY(hipMemcpyAsync(d.vel.get(), d.snap_v.get(), dof_bytes, hipMemcpyDeviceToDevice, d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
        if (cancel_.load()) return result;
        if (substeps >= max_sub) {
            if (crossed) return fail(KHS_ERROR_NOT_CONVERGED, "Animated collider crossed hair after adaptive substepping.");
            return result;
        }
        substeps = std::min(max_sub, substeps * 2);
    }
    if (!animation_ready_) {
        const size_t n3 = 3 * size_t(d.N) * sizeof(double);
        ++d.launch_count;
        HIP_TRY(hipMemcpyAsync(d.root_cur.get(), d.root_pend.get(), n3, hipMemcpyDeviceToDevice, d.stream));
        ++d.launch_count;
        HIP_TRY(hipMemcpyAsync(d.rot_cur.get(), d.rot_pend.get(), n3, hipMemcpyDeviceToDevice, d.stream));
        if (d.T) HIP_TRY(hipMemcpyAsync(d.col_start.get(), d.col_pend.get(), col_bytes, hipMemcpyDeviceToDevice, d.stream));
    }
    clear_failure();
    step_stats_.phase = KHS_PHASE_FINISHED;
    phase_.store(KHS_PHASE_FINISHED);
    HIP_TRY(hipEventRecord(d.frame_end, d.stream));
    HIP_TRY(hipEventSynchronize(d.frame_end));
    d.settle();
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, d.frame_begin, d.frame_end));
    gpu_stats_.last_frame_milliseconds = ms;
    // Phase times from the per-launch HIP events (with kernel timing on):
    // assembly = gradient and line-search evaluations with their list
    // passes, collision = tree rebuilds, sweeps and step bounds,
    // optimization = the block solves.
    if (d.timing) {
        gpu_stats_.last_assembly_milliseconds =
            (d.totals.assembly_ms - before.assembly_ms) + (d.totals.energy_ms - before.energy_ms);
        gpu_stats_.last_collision_milliseconds = d.totals.collision_ms - before.collision_ms;
        gpu_stats_.last_optimization_milliseconds = d.totals.solve_ms - before.solve_ms;
    } else {
        gpu_stats_.last_assembly_milliseconds = gpu_stats_.last_collision_milliseconds =
            gpu_stats_.last_optimization_milliseconds = 0.0;
    }
    d.substep_ms.assign(substeps, 0.0);
    for (uint32_t k = 0; k < substeps; ++k) {
        float sms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&sms, d.substep_events[k], k + 1 < substeps ? d.substep_events[k + 1] : d.frame_end));
        d.substep_ms[k] = sms;
    }
    d.totals.substep_count = substeps;
    error_.clear();
    return KHS_OK;
}

KhsResult Solver::step(double frame_dt)
{
    if (!built_) {
        error_ = "Build the solver first.";
        return KHS_ERROR_INVALID_STATE;
    }
    if (!(frame_dt > 0.0) || !std::isfinite(frame_dt)) {
        error_ = "The time step must be positive and finite.";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    DeviceState &d = *dev_;
    return run_frame(frame_dt, d.root_cur.get(), d.rot_cur.get(), d.root_pend.get(), d.rot_pend.get(),
                     d.T ? d.col_start.get() : nullptr, d.T ? d.col_pend.get() : nullptr, 0, 0);
}

KhsResult Solver::step_animation_frame(uint32_t frame, double frame_dt)
{
    // reference cuda_backend.cu:488-498
    if (!built_) return KHS_ERROR_INVALID_STATE;
    if (!animation_ready_ || frame >= animation_frames_) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid animation frame index.");
    if (frame == 0) {
        current_frame_ = 0;
        keep_in_ring(0);
        return KHS_OK;
    }
    if (frame != current_frame_ + 1) return fail(KHS_ERROR_INVALID_STATE, "Animation frames must be stepped in order.");
    DeviceState &d = *dev_;
    const size_t n3 = 3 * size_t(d.N), v3 = 3 * size_t(d.V);
    const uint32_t a = frame - 1, b = frame;
    const KhsResult r = run_frame(frame_dt, d.root_anim.get() + a * n3, d.rot_anim.get() + a * n3,
                                  d.root_anim.get() + b * n3, d.rot_anim.get() + b * n3,
                                  d.T ? d.col_anim.get() + a * v3 : nullptr, d.T ? d.col_anim.get() + b * v3 : nullptr,
                                  frame, animation_frames_);
    if (r == KHS_OK) {
        current_frame_ = frame;
        keep_in_ring(frame);
    }
    return r;
}

// Keeps the current state as that of `frame` in its rollback slot.
void Solver::keep_in_ring(uint32_t frame)
{
    if (ring_frame_.empty()) return;
    DeviceState &d = *dev_;
    const size_t dofs = 6 * size_t(d.N), slot = frame % ring_frame_.size();
    ++d.launch_count;
    HIP_TRY(hipMemcpyAsync(d.ring_q.get() + slot * dofs, d.q.get(), dofs * sizeof(double), hipMemcpyDeviceToDevice,
                           d.stream));
    ++d.launch_count;
    HIP_TRY(hipMemcpyAsync(d.ring_v.get() + slot * dofs, d.vel.get(), dofs * sizeof(double), hipMemcpyDeviceToDevice,
                           d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    ring_frame_[slot] = frame;
}

// ------------------------------------------------------------ rollback ring
KhsResult Solver::set_rollback_ring(uint32_t frames)
{
    if (!built_) return KHS_ERROR_INVALID_STATE;
    DeviceState &d = *dev_;
    try {
        const size_t dofs = 6 * size_t(d.N);
        d.ring_q.allocate(size_t(frames) * dofs);
        d.ring_v.allocate(size_t(frames) * dofs);
    } catch (const std::exception &e) {
        ring_frame_.clear();
        return fail(KHS_ERROR_OUT_OF_MEMORY, e.what());
    }
    ring_frame_.assign(frames, ~0u);
    gpu_stats_.resident_bytes = d.resident_bytes();
    return KHS_OK;
}

KhsResult Solver::rollback(uint32_t frame)
{
    if (!built_ || !animation_ready_) return KHS_ERROR_INVALID_STATE;
    if (ring_frame_.empty() || ring_frame_[frame % ring_frame_.size()] != frame)
        return fail(KHS_ERROR_INVALID_ARGUMENT, "The rollback ring does not hold that frame.");
    DeviceState &d = *dev_;
    const size_t dofs = 6 * size_t(d.N), slot = frame % ring_frame_.size();
    const size_t bytes = dofs * sizeof(double);
    const double *q = d.ring_q.get() + slot * dofs, *v = d.ring_v.get() + slot * dofs;
    for (double *dst : {d.q.get(), d.q_old.get(), d.snap_q.get()}) {
        ++d.launch_count;
        HIP_TRY(hipMemcpyAsync(dst, q, bytes, hipMemcpyDeviceToDevice, d.stream));
    }
    for (double *dst : {d.vel.get(), d.snap_v.get()}) {
        ++d.launch_count;
        HIP_TRY(hipMemcpyAsync(dst, v, bytes, hipMemcpyDeviceToDevice, d.stream));
    }
    HIP_TRY(hipStreamSynchronize(d.stream));
    restore_state(frame);
    return KHS_OK;
}

// Solver bookkeeping of a restored state (checkpoint or rollback slot).
void Solver::restore_state(uint32_t frame)
{
    current_frame_ = frame;
    cancel_.store(false);
    frame_.store(frame);
    substep_.store(0);
    substep_count_.store(0);
    iteration_.store(0);
    phase_.store(KHS_PHASE_FINISHED);
    step_stats_ = {};
    step_stats_.struct_size = sizeof(step_stats_);
    step_stats_.phase = KHS_PHASE_FINISHED;
    clear_failure();
    error_.clear();
}

// ------------------------------------------------------------ checkpoints
uint64_t Solver::checkpoint_size() const
{
    if (!built_ || !animation_ready_) return 0;
    return 24 + 2 * dof_count() * sizeof(double);
}

KhsResult Solver::save_checkpoint(void *data, uint64_t capacity)
{
    // reference cuda_backend.cu:506-517: {magic, abi, frame, dof} + q + v
    if (!built_) return KHS_ERROR_INVALID_STATE;
    const uint64_t required = checkpoint_size();
    if (!data || !required || capacity < required) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid checkpoint output buffer.");
    unsigned char *out = static_cast<unsigned char *>(data);
    const uint64_t magic = kCheckpointMagic, dof = dof_count();
    const uint32_t abi = KHS_ABI_VERSION, frame = current_frame_;
    std::memcpy(out, &magic, 8);
    std::memcpy(out + 8, &abi, 4);
    std::memcpy(out + 12, &frame, 4);
    std::memcpy(out + 16, &dof, 8);
    dev_->q.download(reinterpret_cast<double *>(out + 24), dof);
    dev_->vel.download(reinterpret_cast<double *>(out + 24 + dof * 8), dof);
    error_.clear();
    return KHS_OK;
}

KhsResult Solver::restore_checkpoint(const void *data, uint64_t size)
{
    // reference cuda_backend.cu:519-537
    if (!built_) return KHS_ERROR_INVALID_STATE;
    const uint64_t required = checkpoint_size();
    if (!data || !required || size != required) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid checkpoint size.");
    const unsigned char *in = static_cast<const unsigned char *>(data);
    uint64_t magic = 0, dof = 0;
    uint32_t abi = 0, frame = 0;
    std::memcpy(&magic, in, 8);
    std::memcpy(&abi, in + 8, 4);
    std::memcpy(&frame, in + 12, 4);
    std::memcpy(&dof, in + 16, 8);
    if (magic != kCheckpointMagic || abi != KHS_ABI_VERSION || dof != dof_count() || frame >= animation_frames_)
        return fail(KHS_ERROR_INVALID_ARGUMENT, "Checkpoint does not match this solver.");
    std::vector<double> q(dof), v(dof);
    std::memcpy(q.data(), in + 24, dof * 8);
    std::memcpy(v.data(), in + 24 + dof * 8, dof * 8);
    DeviceState &d = *dev_;
    d.q.upload(q.data(), dof);
    d.vel.upload(v.data(), dof);
    d.q_old.upload(q.data(), dof);
    d.snap_q.upload(q.data(), dof);
    d.snap_v.upload(v.data(), dof);
    restore_state(frame);
    return KHS_OK;
}

// ----------------------------------------------------------------- output
KhsResult Solver::copy_original_positions(KhsVec3 *positions, uint32_t capacity) const
{
    if (!built_ || !positions || capacity < points_.size()) return KHS_ERROR_INVALID_ARGUMENT;
    try {
        DeviceState &d = *dev_;
        KHS_LAUNCH(d, dev::k_gather_positions, dim3(grid_for(d.P, kBlock)), dim3(kBlock), 0, d.stream, d.q.get(),
                           d.mapping.get(), d.P, d.out3.get());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(d.stream));
        d.out3.download(reinterpret_cast<double *>(positions), 3 * size_t(d.P));
        return KHS_OK;
    } catch (...) {
        return KHS_ERROR_INTERNAL;
    }
}

// Appends the current original-point positions to a .khc cache file: the
// gather runs on the device, one copy lands in pinned staging, and the
// bytes go to the descriptor as they are (3P little-endian doubles, the
// frame record of blender_extension/kami_hair_solver/addon.py:580-584).
KhsResult Solver::write_cache_frame(int fd)
{
    if (!built_) return KHS_ERROR_INVALID_STATE;
    if (fd < 0) return KHS_ERROR_INVALID_ARGUMENT;
    DeviceState &d = *dev_;
    const size_t bytes = 3 * size_t(d.P) * sizeof(double);
    if (!d.cache_host) HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&d.cache_host), bytes, hipHostMallocDefault));
    KHS_LAUNCH(d, dev::k_gather_positions, dim3(grid_for(d.P, kBlock)), dim3(kBlock), 0, d.stream, d.q.get(),
                       d.mapping.get(), d.P, d.out3.get());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(d.cache_host, d.out3.get(), bytes, hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    const char *p = reinterpret_cast<const char *>(d.cache_host);
    for (size_t left = bytes; left;) {
        const ssize_t w = ::write(fd, p, left);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) {
            error_ = std::string("Writing the cache frame failed: ") + std::strerror(w < 0 ? errno : EIO);
            return KHS_ERROR_INTERNAL;
        }
        p += w;
        left -= size_t(w);
    }
    return KHS_OK;
}

KhsResult Solver::copy_internal_positions(KhsVec3 *positions, uint32_t capacity) const
{
    if (!built_ || !positions || capacity < model_.node_count) return KHS_ERROR_INVALID_ARGUMENT;
    try {
        DeviceState &d = *dev_;
        KHS_LAUNCH(d, dev::k_gather_positions, dim3(grid_for(d.N, kBlock)), dim3(kBlock), 0, d.stream, d.q.get(),
                           (const uint32_t *)nullptr, d.N, d.out3.get());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(d.stream));
        d.out3.download(reinterpret_cast<double *>(positions), 3 * size_t(d.N));
        return KHS_OK;
    } catch (...) {
        return KHS_ERROR_INTERNAL;
    }
}

KhsResult Solver::copy_mapping(uint32_t *indices, uint32_t capacity) const
{
    if (!built_ || !indices || capacity < points_.size()) return KHS_ERROR_INVALID_ARGUMENT;
    dev_->mapping.download(indices, points_.size());
    return KHS_OK;
}

KhsResult Solver::build_stats(KhsBuildStats *stats) const
{
    if (!built_ || !stats || stats->struct_size < sizeof(KhsBuildStats)) return KHS_ERROR_INVALID_ARGUMENT;
    *stats = build_stats_;
    return KHS_OK;
}

KhsResult Solver::step_stats(KhsStepStats *stats) const
{
    if (!stats || stats->struct_size < sizeof(KhsStepStats)) return KHS_ERROR_INVALID_ARGUMENT;
    *stats = step_stats_;
    return KHS_OK;
}

KhsResult Solver::gpu_stats(KhsGpuStats *stats) const
{
    if (!built_ || !stats || stats->struct_size < sizeof(KhsGpuStats)) return KHS_ERROR_INVALID_ARGUMENT;
    *stats = gpu_stats_;
    return KHS_OK;
}

KhsResult Solver::progress(KhsProgress *out) const
{
    if (!built_ || !out || out->struct_size < sizeof(KhsProgress)) return KHS_ERROR_INVALID_ARGUMENT;
    KhsProgress p{};
    p.struct_size = sizeof(p);
    p.phase = phase_.load();
    p.frame_index = frame_.load();
    p.frame_count = frame_count_.load();
    p.substep = substep_.load();
    p.substep_count = substep_count_.load();
    p.nonlinear_iteration = iteration_.load();
    p.nonlinear_iteration_limit = desc_.newton_iterations;
    p.cancelled = cancel_.load() ? 1u : 0u;
    const auto start = std::chrono::steady_clock::time_point(std::chrono::steady_clock::duration(frame_clock_.load()));
    p.frame_elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    *out = p;
    return KHS_OK;
}

KhsResult Solver::failure_diagnostics(KhsFailureDiagnostics *diagnostics) const
{
    if (!built_ || !diagnostics || diagnostics->struct_size < sizeof(KhsFailureDiagnostics))
        return KHS_ERROR_INVALID_ARGUMENT;
    *diagnostics = failure_;
    return KHS_OK;
}

KhsResult Solver::request_cancel()
{
    if (!built_) return KHS_ERROR_INVALID_STATE;
    cancel_.store(true);
    return KHS_OK;
}

// ------------------------------------------------------------- test hooks
KhsResult Solver::get_state(double *q, double *v, uint64_t count) const
{
    if (!built_ || count != dof_count()) return KHS_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipStreamSynchronize(dev_->stream));
    if (q) dev_->q.download(q, count);
    if (v) dev_->vel.download(v, count);
    return KHS_OK;
}

KhsResult Solver::set_state(const double *q, const double *q_old, const double *v, uint64_t count)
{
    if (!built_ || !q || !v || count != dof_count()) return KHS_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipStreamSynchronize(dev_->stream));
    dev_->q.upload(q, count);
    dev_->vel.upload(v, count);
    dev_->q_old.upload(q_old ? q_old : q, count);
    return KHS_OK;
}

KhsResult Solver::copy_node_model(double *frames9, double *mass, double *rotational_mass, uint32_t *flags,
                                  uint32_t capacity) const
{
    if (!built_ || capacity < model_.node_count) return KHS_ERROR_INVALID_ARGUMENT;
    const DeviceState &d = *dev_;
    if (frames9) d.frame.download(frames9, 9 * size_t(d.N));
    if (mass) d.mass.download(mass, d.N);
    if (rotational_mass) d.rmass.download(rotational_mass, d.N);
    if (flags) {
        std::vector<uint32_t> f(d.N);
        d.flags.download(f.data(), d.N);
        for (uint32_t n = 0; n < d.N; ++n)
            flags[n] = ((f[n] & kNodeFixed) ? 1u : 0u) | ((f[n] & kNodeVirtual) ? 2u : 0u);
    }
    return KHS_OK;
}

KhsResult Solver::copy_element_model(uint32_t *nodes2, uint32_t *strand, uint32_t *contact, double *rest7,
                                     uint32_t capacity) const
{
    if (!built_ || capacity < model_.element_count) return KHS_ERROR_INVALID_ARGUMENT;
    const DeviceState &d = *dev_;
    std::vector<double> rest(8 * size_t(d.N));
    d.rest.download(rest.data(), rest.size());
    std::vector<uint32_t> f(d.N), so(d.N);
    d.flags.download(f.data(), d.N);
    d.strand_of.download(so.data(), d.N);
    uint32_t e = 0;
    for (uint32_t n = 0; n < d.N; ++n) {
        if (!(f[n] & kNodeHasElement)) continue;
        if (nodes2) { nodes2[2 * e] = n; nodes2[2 * e + 1] = n + 1; }
        if (strand) strand[e] = so[n];
        if (contact) contact[e] = (f[n] & kNodeContact) ? 1u : 0u;
        if (rest7) for (int k = 0; k < 7; ++k) rest7[7 * size_t(e) + k] = rest[8 * size_t(n) + k];
        ++e;
    }
    return KHS_OK;
}

KhsResult Solver::test_evaluate(double dt, double *gradient, double *direction, KhsxEvaluation *out)
{
    if (!built_ || !(dt > 0.0) || !out || out->struct_size < sizeof(KhsxEvaluation)) return KHS_ERROR_INVALID_ARGUMENT;
    DeviceState &d = *dev_;
    if (d.T) rebuild_bvh(false);
    const Reduced r = evaluate(true, dt, false);
    KhsxEvaluation e{};
    e.struct_size = sizeof(e);
    e.feasible = r.feasible ? 1u : 0u;
    e.candidate_pairs = r.candidates;
    e.active_contacts = r.active;
    e.objective = r.objective;
    e.elastic = r.elastic;
    e.contact = r.contact;
    e.kinetic = r.kinetic;
    e.minimum_gap = r.min_gap;
    if (gradient) d.grad.download(gradient, dof_count());
    if (direction && direction_mode_ == 1) {
        // the forced fallback: -P g, no trust-radius scaling
        const RodView rv = d.rods();
        launch_preconditioner(dt);
        KHS_LAUNCH(d, dev::k_force_steepest, dim3(1), dim3(1), 0, d.stream, d.choice(), d.mirror_choice());
        KHS_LAUNCH(d, dev::k_steepest, dim3(grid_for(6 * size_t(d.N), kBlock)), dim3(kBlock), 0, d.stream, rv,
                           d.choice(), d.grad.get(), d.precond.get(), d.dir.get());
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(d.stream));
        d.dir.download(direction, dof_count());
    } else if (direction) {
        const RodView rv = d.rods();
        KHS_LAUNCH(d, dev::k_fill_u32, dim3(1), dim3(64), 0, d.stream, d.flag.get(), size_t(1), 0u);
        launch_block_solve(d.flag.get(), nullptr, nullptr);
        unsigned failed = 0;
        HIP_TRY(hipMemcpyAsync(&failed, d.flag.get(), sizeof(unsigned), hipMemcpyDeviceToHost, d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
        e.solve_failed = failed;
        d.dir.download(direction, dof_count());
    }
    *out = e;
    return KHS_OK;
}

KhsResult Solver::broadphase_pairs(uint32_t *pairs, uint64_t capacity, uint64_t *count)
{
    if (!built_ || !count) return KHS_ERROR_INVALID_ARGUMENT;
    DeviceState &d = *dev_;
    if (!d.T) {
        *count = 0;
        return KHS_OK;
    }
    rebuild_bvh(false);
    Buffer<uint32_t> out;
    out.allocate(std::max<uint64_t>(2 * capacity, 2));
    const StepParams p = step_params(1.0);
    KHS_LAUNCH(d, dev::k_fill_u64, dim3(1), dim3(64), 0, d.stream, d.key64.get(), size_t(1), 0ull);
    KHS_LAUNCH(d, dev::k_broadphase_pairs, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.rods(),
                       d.collider(), p, d.q.get(), out.get(), (unsigned long long)capacity, d.key64.get());
    HIP_TRY(hipGetLastError());
    unsigned long long total = 0;
    HIP_TRY(hipMemcpyAsync(&total, d.key64.get(), sizeof(total), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    *count = total;
    if (pairs && capacity) {
        const uint64_t n = std::min<uint64_t>(total, capacity);
        std::vector<uint64_t> keyed(n);
        std::vector<uint32_t> raw(2 * n);
        out.download(raw.data(), 2 * n);
        for (uint64_t i = 0; i < n; ++i) keyed[i] = (uint64_t(raw[2 * i]) << 32) | raw[2 * i + 1];
        std::sort(keyed.begin(), keyed.end());
        for (uint64_t i = 0; i < n; ++i) {
            pairs[2 * i] = uint32_t(keyed[i] >> 32);
            pairs[2 * i + 1] = uint32_t(keyed[i]);
        }
    }
    return (pairs && total > capacity) ? KHS_ERROR_INVALID_ARGUMENT : KHS_OK;
}

// khsxBroadphasePairs without any allocation: the pairs are staged in the
// file-scope array above (its counter cleared by a fill kernel) and sorted
// in the caller's buffer. Overlapping calls on one GPU share the array and
// race.
KhsResult Solver::broadphase_pairs_no_alloc(uint32_t *pairs, uint64_t capacity, uint64_t *count)
{
    if (!built_ || !count) return KHS_ERROR_INVALID_ARGUMENT;
    if (capacity > kNoAllocPairs) {
        error_ = "khsxBroadphasePairsNoAlloc stages at most " + std::to_string(kNoAllocPairs) + " pairs; capacity " +
                 std::to_string(capacity) + " is larger.";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    DeviceState &d = *dev_;
    if (!d.T) {
        *count = 0;
        return KHS_OK;
    }
    uint32_t *staging = no_alloc_staging();
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(staging);
    uint32_t *staged = staging + kNoAllocHeader;
    rebuild_bvh(false);
    const StepParams p = step_params(1.0);
    KHS_LAUNCH(d, dev::k_fill_u64, dim3(1), dim3(64), 0, d.stream, counter, size_t(1), 0ull);
    KHS_LAUNCH(d, dev::k_broadphase_pairs, dim3(d.node_blocks()), dim3(kBlock), 0, d.stream, d.rods(),
                       d.collider(), p, d.q.get(), staged, (unsigned long long)capacity, counter);
    HIP_TRY(hipGetLastError());
    unsigned long long total = 0;
    HIP_TRY(hipMemcpyAsync(&total, counter, sizeof(total), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    *count = total;
    if (pairs && capacity) {
        const uint64_t n = std::min<uint64_t>(total, capacity);
        HIP_TRY(hipMemcpyAsync(pairs, staged, n * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
        struct Pair {
            uint32_t element, triangle;
            bool operator<(const Pair &o) const
            {
                return element != o.element ? element < o.element : triangle < o.triangle;
            }
        };
        static_assert(sizeof(Pair) == 2 * sizeof(uint32_t) && alignof(Pair) == alignof(uint32_t), "pair layout");
        Pair *first = reinterpret_cast<Pair *>(pairs);
        std::sort(first, first + n);
    }
    return (pairs && total > capacity) ? KHS_ERROR_INVALID_ARGUMENT : KHS_OK;
}

KhsResult Solver::copy_blocks(double *diagonal, double *upper, uint32_t capacity) const
{
    if (!built_ || capacity < model_.node_count) return KHS_ERROR_INVALID_ARGUMENT;
    HIP_TRY(hipStreamSynchronize(dev_->stream));
    if (diagonal) dev_->Dblk.download(diagonal, 36 * size_t(dev_->N));
    if (upper) dev_->Ublk.download(upper, 36 * size_t(dev_->N));
    return KHS_OK;
}

KhsResult Solver::check_bvh(uint64_t *violations)
{
    if (!built_ || !violations) return KHS_ERROR_INVALID_ARGUMENT;
    DeviceState &d = *dev_;
    for (int k = 0; k < 6; ++k) violations[k] = 0;
    if (d.T < 2) return KHS_OK;
    rebuild_bvh(false);
    Buffer<unsigned long long> out;
    out.allocate(6);
    KHS_LAUNCH(d, dev::k_fill_u64, dim3(1), dim3(64), 0, d.stream, out.get(), size_t(6), 0ull);
    KHS_LAUNCH(d, dev::k_check_bvh, dim3(grid_for(d.T, kBlock)), dim3(kBlock), 0, d.stream, d.T,
                       d.keys_sorted.get(), d.ids_sorted.get(), d.tri_box.get(), d.parent_internal.get(),
                       d.parent_leaf.get(), d.bvh.get(), out.get());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(d.stream));
    unsigned long long h[6];
    out.download(h, 6);
    for (int k = 0; k < 6; ++k) violations[k] = h[k];
    return KHS_OK;
}

KhsResult Solver::set_convergence_mode(uint32_t mode)
{
    if (mode > 1) return KHS_ERROR_INVALID_ARGUMENT;
    convergence_mode_ = mode;
    return KHS_OK;
}

KhsResult Solver::set_precision_mode(uint32_t mode)
{
    if (mode > 1) {
        error_ = "Precision mode must be 0 (FP64) or 1 (FP32 block solve).";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    precision_mode_ = mode;
    return KHS_OK;
}

KhsResult Solver::set_contact_reuse(uint32_t mode)
{
    if (mode > 1) {
        error_ = "Contact reuse must be 0 (off) or 1 (on).";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    contact_reuse_ = mode == 1;
    return KHS_OK;
}

KhsResult Solver::set_solve_algorithm(uint32_t algorithm)
{
    if (algorithm > 2) {
        error_ = "Solve algorithm must be 0 (block Thomas), 1 (twisted) or 2 (fused).";
        return KHS_ERROR_INVALID_ARGUMENT;
    }
    solve_algorithm_ = algorithm;
    return KHS_OK;
}

KhsResult Solver::set_kernel_timing(uint32_t enable)
{
    if (!built_) return KHS_ERROR_INVALID_STATE;
    dev_->timing = enable != 0;
    dev_->totals = KhsxKernelTiming{};
    return KHS_OK;
}

KhsResult Solver::kernel_timing(KhsxKernelTiming *out) const
{
    if (!built_ || !out || out->struct_size < sizeof(KhsxKernelTiming)) return KHS_ERROR_INVALID_ARGUMENT;
    KhsxKernelTiming t = dev_->totals;
    t.struct_size = sizeof(t);
    t.nodes = model_.node_count;
    *out = t;
    return KHS_OK;
}

KhsResult Solver::substep_times(double *ms, uint32_t capacity, uint32_t *count) const
{
    if (!built_ || !count) return KHS_ERROR_INVALID_ARGUMENT;
    const std::vector<double> &v = dev_->substep_ms;
    *count = uint32_t(v.size());
    if (ms)
        for (uint32_t i = 0; i < std::min<uint32_t>(capacity, uint32_t(v.size())); ++i) ms[i] = v[i];
    return KHS_OK;
}

KhsResult Solver::contact_list_stats(KhsxContactListStats *out)
{
    if (!out || out->struct_size < sizeof(KhsxContactListStats)) return KHS_ERROR_INVALID_ARGUMENT;
    if (!built_) return fail(KHS_ERROR_INVALID_STATE, "The solver is not built.");
    DeviceState &d = *dev_;
    KhsxContactListStats r{};
    r.struct_size = sizeof(r);
    r.capacity = kCacheCapacity;
    r.elements = d.N;
    std::vector<uint32_t> count(d.N);
    uint32_t listed = 0;
    HIP_TRY(hipMemcpyAsync(count.data(), d.cache_count.get(), d.cache_count.bytes(), hipMemcpyDeviceToHost, d.stream));
    std::vector<uint32_t> flags(d.N);
    HIP_TRY(hipMemcpyAsync(flags.data(), d.flags.get(), d.flags.bytes(), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipMemcpyAsync(&listed, d.cache_list_count.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    r.listed = listed;
    for (uint32_t e = 0; e < d.N; ++e) {
        const uint32_t c = count[e];
        if (c == 0) continue;
        ++r.nonempty;
        const bool root = flags[e] & kNodeFixed;
        r.root_nonempty += root ? 1 : 0;
        if (c == kCacheOverflow) {
            ++r.overflowed;
            continue;
        }
        r.entries += c;
        r.root_entries += root ? c : 0;
        r.longest = std::max<uint64_t>(r.longest, c);
        int bin = 0;
        while (bin < 8 && (2u << bin) <= c) ++bin;
        ++r.length_histogram[bin];
    }
    *out = r;
    return KHS_OK;
}

double Solver::max_radius() const
{
    return strand_r_.empty() ? material_.radius : strand_r_max_;
}

// Uploads the per-strand materials of a built solver: (E, r) of each node's
// strand for the step's kernels, the strands' mass densities, and the
// masses and trust radius they give.
void Solver::apply_strand_materials()
{
    DeviceState &d = *dev_;
    if (strand_r_.empty()) {
        d.elem_material.reset();
        d.strand_density.reset();
    } else {
        std::vector<uint32_t> strand_of(d.N);
        d.strand_of.download(strand_of.data(), d.N);
        std::vector<double> per_node(2 * size_t(d.N)), density(2 * size_t(d.S));
        for (uint32_t n = 0; n < d.N; ++n) {
            per_node[2 * size_t(n)] = strand_E_[strand_of[n]];
            per_node[2 * size_t(n) + 1] = strand_r_[strand_of[n]];
        }
        for (uint32_t k = 0; k < d.S; ++k) {
            KhsHairMaterial m = material_;
            m.radius = strand_r_[k];
            const MassDensity rho = mass_density(m);
            density[2 * size_t(k)] = rho.translational;
            density[2 * size_t(k) + 1] = rho.rotational;
        }
        d.elem_material.allocate(per_node.size());
        d.elem_material.upload(per_node.data(), per_node.size());
        d.strand_density.allocate(density.size());
        d.strand_density.upload(density.data(), density.size());
    }
    const MassDensity rho = mass_density(material_);
    KHS_LAUNCH(d, dev::k_model_masses, dim3(grid_for(d.S, kBlock)), dim3(kBlock), 0, d.stream, d.rods(),
                       (const double *)d.rest.get(), rho.translational, rho.rotational,
                       d.strand_density.size() ? (const double *)d.strand_density.get() : nullptr, d.mass.get(),
                       d.rmass.get());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(d.stream));
    d.cache_valid = false;
    trust_radius_ = std::max(0.25 * model_.maximum_rest_length,
                             8.0 * (max_radius() + material_.collider_offset + material_.barrier_distance));
}

KhsResult Solver::set_strand_materials(const double *E, const double *r, uint32_t count)
{
    if (!E && !r) {
        strand_E_.clear();
        strand_r_.clear();
    } else {
        if (!E || !r || !has_hair_ || count + 1 != offsets_.size()) {
            error_ = "Per-strand materials need both arrays, one value per hair strand.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        for (uint32_t k = 0; k < count; ++k)
            if (!(E[k] > 0.0 && std::isfinite(E[k]) && r[k] > 0.0 && std::isfinite(r[k]))) {
                error_ = "Strand " + std::to_string(k) + " has a material value out of range.";
                return KHS_ERROR_INVALID_ARGUMENT;
            }
        strand_E_.assign(E, E + count);
        strand_r_.assign(r, r + count);
        strand_r_max_ = *std::max_element(strand_r_.begin(), strand_r_.end());
    }
    if (built_) apply_strand_materials();
    return KHS_OK;
}

KhsResult Solver::set_collider_state(const double *current, const double *previous, uint32_t count)
{
    if (!built_ || !dev_->T || !current || count != dev_->V) return KHS_ERROR_INVALID_ARGUMENT;
    DeviceState &d = *dev_;
    HIP_TRY(hipStreamSynchronize(d.stream));
    d.col_cur.upload(current, 3 * size_t(count));
    d.col_old.upload(previous ? previous : current, 3 * size_t(count));
    d.cache_valid = false;
    return KHS_OK;
}

// The friction stage as a substep ends: the tree swept over the collider's
// motion and every list gathered, as at substep start, then the list pass
// and the impulse.
KhsResult Solver::friction_stage(double dt)
{
    if (!built_ || !(dt > 0.0) || !std::isfinite(dt)) return KHS_ERROR_INVALID_ARGUMENT;
    DeviceState &d = *dev_;
    if (d.T) {
        rebuild_bvh(true);
        refresh_contact_cache(dev::kAtState, true, true);
        if (material_.friction > 0.0) enqueue_friction(step_params(dt));
    }
    HIP_TRY(hipStreamSynchronize(d.stream));
    return KHS_OK;
}

// The line search's step bound along a given direction, by the step's own
// path: lists gathered at the state, then the swept check and the bound
// kernel (lists, new gathers and overflowed lists).
KhsResult Solver::step_bound(const double *direction, uint64_t count, double *alpha)
{
    if (!built_ || !direction || !alpha || count != dof_count()) return KHS_ERROR_INVALID_ARGUMENT;
    DeviceState &d = *dev_;
    HIP_TRY(hipStreamSynchronize(d.stream));
    d.dir.upload(direction, count);
    if (!d.T) {
        *alpha = 1.0;
        return KHS_OK;
    }
    rebuild_bvh(true);
    refresh_contact_cache(dev::kAtState, true, true);
    const StepParams cp = step_params(1.0);
    swept_bound_async(cp, nullptr);
    d.reduce_async(d.node_blocks(), DeviceState::kCcdSlot, dev::ConvergenceTest{0.0, 0.0, 0.0, nullptr},
                   dev::ListPartials{nullptr, d.list_alpha.get(), kSweptBoundBlocks});
    *alpha = d.read_back().ccd.alpha;
    return KHS_OK;
}

KhsResult Solver::sweep_check(uint32_t *feasible, KhsFailureDiagnostics *diagnostics)
{
    if (!built_ || !feasible) return KHS_ERROR_INVALID_ARGUMENT;
    if (diagnostics && diagnostics->struct_size < sizeof(KhsFailureDiagnostics)) return KHS_ERROR_INVALID_ARGUMENT;
    DeviceState &d = *dev_;
    SweepRecord rec{};
    if (d.T) {
        enqueue_sweep(step_params(1.0));
        HIP_TRY(hipMemcpyAsync(&rec, d.sweep.get(), sizeof(rec), hipMemcpyDeviceToHost, d.stream));
    }
    d.wait();
    *feasible = rec.claimed ? 0u : 1u;
    if (diagnostics) {
        KhsFailureDiagnostics f{};
        f.struct_size = sizeof(f);
        f.strand_index = f.element_index = f.collider_triangle_index = 0xffffffffu;
        if (rec.claimed) {
            f.kind = KHS_FAILURE_MOVING_COLLIDER_SWEEP;
            f.strand_index = rec.strand;
            f.element_index = rec.element;
            f.collider_triangle_index = rec.triangle;
            f.distance = rec.distance;
            f.required_distance = rec.required;
            f.clearance = rec.clearance;
            f.collider_substep_displacement = rec.displacement;
            f.hair_start = {rec.hair_start[0], rec.hair_start[1], rec.hair_start[2]};
            f.hair_end = {rec.hair_end[0], rec.hair_end[1], rec.hair_end[2]};
            f.collider_point = {rec.collider_point[0], rec.collider_point[1], rec.collider_point[2]};
        }
        *diagnostics = f;
    }
    return KHS_OK;
}

KhsResult Solver::set_direction_mode(uint32_t mode)
{
    if (mode > 1) return KHS_ERROR_INVALID_ARGUMENT;
    direction_mode_ = mode;
    return KHS_OK;
}

KhsResult Solver::direction_counts(uint64_t *block, uint64_t *steepest) const
{
    if (!block || !steepest) return KHS_ERROR_INVALID_ARGUMENT;
    *block = block_directions_;
    *steepest = steepest_directions_;
    return KHS_OK;
}

KhsResult Solver::launch_counts(uint64_t *dispatches, uint64_t *host_waits) const
{
    if (!dispatches || !host_waits) return KHS_ERROR_INVALID_ARGUMENT;
    *dispatches = built_ ? dev_->launch_count : 0;
    *host_waits = built_ ? dev_->wait_count : 0;
    return KHS_OK;
}

bool query_device(KhsGpuInfo &info, std::string &error)
{
    std::memset(&info, 0, sizeof(info));
    info.struct_size = sizeof(info);
    int count = 0;
    const hipError_t r = hipGetDeviceCount(&count);
    if (r != hipSuccess || count < 1) {
        info.device_ordinal = -1;
        error = r == hipSuccess ? "No HIP GPU found." : hipGetErrorString(r);
        return false;
    }
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        error = "Failed to query the HIP device.";
        return false;
    }
    info.available = 1;
    info.device_ordinal = 0;
    // gfx950 reports major 9, minor 5
    info.compute_capability_major = uint32_t(prop.major);
    info.compute_capability_minor = uint32_t(prop.minor);
    info.total_vram_bytes = prop.totalGlobalMem;
    std::snprintf(info.device_name, sizeof(info.device_name), "%s (%s)", prop.name, prop.gcnArchName);
    return true;
}

}  // namespace khs
