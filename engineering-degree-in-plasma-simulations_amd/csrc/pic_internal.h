// pic_internal.h -- shared definitions of the gfx950 PIC/DSMC library.
//
// Device data layout (HBM):
//   node fields      flat, i-fastest, u = (k*nj + j)*ni + i
//   phi, rho         double[nn]
//   E                float4[nn] {Ex, Ey, Ez, 0}: one 16-byte load per CIC corner
//   node type        uint8[nn]  (NT_* below; matrix-free operator rows)
//   particles (SoA)  cell  uint32 packed (i | j<<11 | k<<22), DEAD when removed
//                    fx, fy, fz float in [0,1): offset inside the cell, in cells
//                    vx, vy, vz float [m/s], w float (macro weight)
//   Storing the cell and the in-cell offset gives the same sub-cell precision
//   anywhere in the box, makes the logical coordinate (World::XtoL,
//   World.cpp:123-127) free, and the cell word is the counting-sort key.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pic_abi.h"

namespace pic {

constexpr double EPS0 = 8.85418782e-12;  // all.h:15
constexpr double QE = 1.602176565e-19;   // all.h:16
constexpr double ME = 9.10938215e-31;    // all.h:18
constexpr double KB = 1.380648e-23;      // all.h:19
constexpr double PI_REF = 3.141592653;   // all.h:20 (truncated, kept for parity)
constexpr double TWO_PI = 6.283185307179586;

constexpr uint32_t DEAD = 0xFFFFFFFFu;
constexpr int MAX_RECT = 8;
constexpr int MAX_SPECIES = 8;

// node types of the matrix-free operator (PotentialSolver.cpp:414-471)
enum : uint8_t {
    NT_REG = 0,
    NT_NEU_IM = 1,  // i == 0     -> neighbour +1
    NT_NEU_IP = 2,  // i == ni-1  -> neighbour -1
    NT_NEU_JM = 3,
    NT_NEU_JP = 4,
    NT_NEU_KM = 5,
    NT_NEU_KP = 6,
    NT_DIR = 7,
};

struct GridC {
    int ni, nj, nk;   // nodes
    int ci, cj, ck;   // cells
    int ncells;
    long long nn;
    double x0[3], dx[3], inv_dx[3];
    int nrect;
    // rectangles in logical coordinates (FP64: containment is decided as the
    // reference does, in FP64, Object.cpp:231-238)
    double rlo[MAX_RECT][3], rhi[MAX_RECT][3];
};

__host__ __device__ inline uint32_t pack_cell(int i, int j, int k) {
    return (uint32_t)i | ((uint32_t)j << 11) | ((uint32_t)k << 22);
}
__host__ __device__ inline void unpack_cell(uint32_t c, int& i, int& j, int& k) {
    i = (int)(c & 0x7FFu);
    j = (int)((c >> 11) & 0x7FFu);
    k = (int)(c >> 22);
}
__host__ __device__ inline int cell_index(const GridC& g, int i, int j, int k) {
    return (k * g.cj + j) * g.ci + i;
}
// 1 / node volume (World.cpp:353-367): half volume per boundary axis
__host__ __device__ inline double inv_node_vol(const GridC& g, int i, int j, int k) {
    double s = g.inv_dx[0] * g.inv_dx[1] * g.inv_dx[2];
    if (i == 0 || i == g.ni - 1) s *= 2.0;
    if (j == 0 || j == g.nj - 1) s *= 2.0;
    if (k == 0 || k == g.nk - 1) s *= 2.0;
    return s;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based
__host__ __device__ inline void philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                           uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}
__host__ __device__ inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
// draw `slot` of the stream addressed by (a, b): block slot/2, half slot%2
struct Draws {
    uint32_t a, b, k0, k1, tag;
    uint32_t cache[4];
    int cached_block;
    __host__ __device__ void init(uint32_t a_, uint32_t b_, uint32_t k0_, uint32_t k1_,
                                  uint32_t tag_) {
        a = a_;
        b = b_;
        k0 = k0_;
        k1 = k1_;
        tag = tag_;
        cached_block = -1;
    }
    __host__ __device__ double operator()(int slot) {
        int blk = slot >> 1;
        if (blk != cached_block) {
            philox4x32(a, b, (uint32_t)blk, tag, k0, k1, cache);
            cached_block = blk;
        }
        int h = slot & 1;
        return u53(cache[2 * h], cache[2 * h + 1]);
    }
};

constexpr uint32_t TAG_LOAD = 0x4C4F4144u;  // stream tags
constexpr uint32_t TAG_MCC = 0x4D434301u;
constexpr uint32_t TAG_SURF = 0x53524601u;
constexpr uint32_t TAG_DSMC = 0x44534D01u;
constexpr uint32_t TAG_SRC = 0x53524301u;
constexpr uint32_t TAG_MERGE = 0x4D524701u;

// one species: SoA particle arrays (two buffer sets for the sort) + grids
struct SpeciesBuf {
    uint32_t* cell;
    float* f[3];
    float* v[3];
    float* w;
};

struct Species {
    double mass = 0, charge = 0, mpw0 = 0;
    size_t cap = 0;
    SpeciesBuf buf[2] = {};
    int cur = 0;
    void* block[2] = {nullptr, nullptr};
    uint32_t* rank = nullptr;    // rank of each particle inside its cell bucket
    uint32_t* count = nullptr;   // ncells + 1 buckets (last = removed)
    uint32_t* off = nullptr;     // ncells + 2 exclusive offsets
    uint32_t* d_n = nullptr;     // [0] live count, [1] count the histogram covers, [2] overflow
    size_t n_upper = 0;          // host upper bound of the live count
    bool hist_valid = false;     // rank/count match the current arrays
    bool sorted = false;         // arrays ordered by cell, off valid
    // DSMC births are appended at the tail of a sorted species; the next
    // DSMC step may index only them (pic_sort_tail) while nothing else has
    // changed the particles (tail_ok).  two_part: the arrays are the sorted
    // part [0, bulk_n) with offsets off, then an appended part whose
    // positions tidx lists in cell order, with offsets toff into tidx; the
    // index covers [bulk_n, t_end).  boff, toff_n, tidx_n: the next
    // extension of the index.
    bool tail_ok = false;
    bool two_part = false;
    size_t bulk_n = 0, t_end = 0;
    uint32_t* toff = nullptr;
    uint32_t* tidx = nullptr;
    uint32_t *boff = nullptr, *toff_n = nullptr, *tidx_n = nullptr;
    // particles in random order (a load or an upload): the next sort
    // partitions them by coarse bucket first
    bool scrambled = false;
    double* den = nullptr;       // number density
    double* mom = nullptr;       // n, nvx, nvy, nvz, nuu, nvv, nww (7 * nn)
    uint32_t seq = 0;            // loader stream id
    // behaviour at objects (PIC_SURF_*), the species that receives emitted
    // particles, and the key of the surface draws (seed, push count)
    int surf_mode = PIC_SURF_ABSORB;
    int surf_target = -1;
    uint32_t surf_seed = 0, surf_step = 0;
    uint32_t merge_calls = 0;    // key word of the merge draws
};

struct Mcc {
    int sn, si, se;
    pic_mcc_desc desc;
    double* sig = nullptr;  // nsig energies followed by nsig values
    int nsig = 0;
    double* d_wsv = nullptr;  // [0] current max, [1] step max (as bits for atomicMax)
    unsigned long long* d_stats = nullptr;  // collisions, ionizations, groups, overflow
    // per-workgroup partials of d_stats and of the step maximum (as bits),
    // striped so that the workgroups' atomics spread over NSTRIPE lines
    unsigned long long* d_part = nullptr;
    uint32_t* d_spill = nullptr;  // pick-pool entries past the LDS list
    double wsv_host = 0;
};

// DSMC_MEX between one species or two of equal weight
struct Dsmc {
    int s1, s2;
    double* d_sv = nullptr;                 // [0] sigma v_rel max, [1] the step's maximum
    unsigned long long* d_part = nullptr;   // striped {collisions, groups, step max bits}
    unsigned long long* d_stats = nullptr;  // summed
};

// a cold or warm beam source through one face
struct Src {
    int sid, face;
    double v_drift, den, T, dt;
    double num_micro;  // particles per call (den v_drift A dt)
    uint32_t calls = 0;
};

struct Prof {
    bool on = false;
    struct Rec {
        int cat;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double ms[PIC_K_COUNT] = {};
    long long n[PIC_K_COUNT] = {};
};

}  // namespace pic

struct pic_ctx {
    pic::GridC g;
    hipStream_t stream = nullptr;
    std::string err;
    double dt = 1e-4;  // World.h:25 default
    // host mirrors of the geometry flags (i-fastest)
    std::vector<int32_t> object_id, node_type_in;
    std::vector<double> phi_fixed;
    struct RectH {
        double c[3], half[3], phi;
    };
    std::vector<RectH> rects;
    bool ntype_dirty = true;
    bool has_neumann = true;  // any Neumann face node (set with the node types)
    // device fields
    double* phi = nullptr;
    double* phi_prev = nullptr;  // previous solution (PIC_SOLVE_EXTRAPOLATE), allocated on demand
    void* sor_split = nullptr;  // the SOR sweeps' red-black split phi, rho, node types (on demand)
    size_t sor_split_bytes = 0;
    bool phi_prev_valid = false;
    double* rho = nullptr;
    float4* ef = nullptr;
    // max over the nodes of |Ex|, |Ey|, |Ez| of ef (as float bits, for
    // atomicMax), kept with every write of ef
    float* emax = nullptr;
    uint8_t* ntype = nullptr;
    double* scratch = nullptr;   // reductions and solver scalars
    uint32_t* scan_tmp = nullptr; // block sums of the offset scan
    uint32_t* part = nullptr;     // partition counts, offsets and scan sums of a scrambled sort
    size_t part_words = 0;
    double* work = nullptr;      // solver vectors (allocated on demand)
    size_t work_bytes = 0;
    std::vector<pic::Species> sp;
    std::vector<pic::Mcc> mcc;
    std::vector<pic::Dsmc> dsmc;
    std::vector<pic::Src> src;
    pic::Prof prof;
    void* unet = nullptr;  // learned preconditioner state
    float* dep_corner = nullptr;  // deposits' corner sums per cell (8, or 7 x 8 for the moments; on demand)
    size_t dep_corner_floats = 0;
    int learned_last_its = 0;  // iterations of the last learned solve (first check near it)
    double* pin = nullptr;     // pinned host memory for the solvers' convergence reads
    size_t pin_bytes = 0;
    hipEvent_t solve_ev[2] = {nullptr, nullptr};  // pic_poisson's timing events
};

// ---- host helpers shared between translation units ----
int pic_fail(pic_ctx* ctx, int code, const std::string& msg);
int pic_hip_check(pic_ctx* ctx, hipError_t e, const char* what);
#define PIC_HIP(ctx, call)                                        \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return pic_hip_check(ctx, e_, #call); \
    } while (0)
// Wait for a stream by polling it: the host resumes as soon as the queued
// work ends.  For the solvers' short, frequent convergence checks, where a
// blocking wait left the GPU idle for tens of microseconds per check.
inline hipError_t pic_spin_wait(hipStream_t s) {
    hipError_t e;
    while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
    }
    return e;
}
inline hipError_t pic_spin_event(hipEvent_t ev) {
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
    }
    return e;
}
// profiling brackets around a launch
void pic_prof_begin(pic_ctx* ctx, int cat, hipEvent_t* a, hipEvent_t* b);
void pic_prof_end(pic_ctx* ctx, int cat, hipEvent_t a, hipEvent_t b);
struct PicProfScope {
    pic_ctx* ctx;
    int cat;
    hipEvent_t a = nullptr, b = nullptr;
    PicProfScope(pic_ctx* c, int k) : ctx(c), cat(k) { pic_prof_begin(ctx, cat, &a, &b); }
    ~PicProfScope() { pic_prof_end(ctx, cat, a, b); }
};
int pic_check_sid(pic_ctx* ctx, int sid);
int pic_ensure_ntype(pic_ctx* ctx);
int pic_ensure_work(pic_ctx* ctx, size_t bytes);
int pic_sort_tail(pic_ctx* ctx, int sid);
int pic_join_tail(pic_ctx* ctx, int sid);
// ctx->pin holds at least `bytes` of pinned host memory (device-to-host
// copies into it are asynchronous, unlike copies into pageable memory)
int pic_ensure_pin(pic_ctx* ctx, size_t bytes);
// histogram + ranks of the current arrays (run-aggregated atomics)
int pic_histogram(pic_ctx* ctx, int sid);
// The workgroup's maxima of three non-negative floats into m[0..2] (as
// unsigned bits, which order like the floats; a NaN orders above every
// number).  Every thread of the workgroup (blockDim.x a multiple of 64, at
// most 1024) calls this.
__device__ inline void block_max3_to(float a0, float a1, float a2, float* m) {
    __shared__ float red[16][3];
    for (int o = 32; o > 0; o >>= 1) {
        a0 = fmaxf(a0, __shfl_xor(a0, o, 64));
        a1 = fmaxf(a1, __shfl_xor(a1, o, 64));
        a2 = fmaxf(a2, __shfl_xor(a2, o, 64));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[wv][0] = a0;
        red[wv][1] = a1;
        red[wv][2] = a2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        float v = 0.f;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) v = fmaxf(v, red[w][threadIdx.x]);
        // an atomic only when the value raises the maximum: thousands of
        // workgroups contend for the three words
        unsigned* mp = reinterpret_cast<unsigned*>(m) + threadIdx.x;
        const unsigned bits = __float_as_uint(v);
        if (bits > __hip_atomic_load(mp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(mp, bits);
    }
}

// number density of a sorted species with its per-cell corner sums in
// `corner` (8 * ncells floats): launches only
void pic_density_launch(pic_ctx* ctx, int sid, float* corner);
inline unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g ? g : 1);
}

// XCD-contiguous chunks for grid sweeps: workgroups are dealt round-robin
// over the 8 XCDs, so workgroup b takes logical chunk (b % 8) * (nb / 8) + b / 8
// and each XCD walks one contiguous slab (its L2 then serves the +-k
// neighbour planes).  The grid size must be a multiple of 8 (xcd_blocks).
constexpr int XCD_COUNT = 8;
inline unsigned xcd_blocks(long long items, int block) {
    // about two items per thread: enough workgroups in flight to cover HBM latency
    long long b = (items + block * 2 - 1) / (block * 2);
    if (b > 16384) b = 16384;
    b = (b + XCD_COUNT - 1) / XCD_COUNT * XCD_COUNT;
    return (unsigned)b;
}
// one item per thread (latency-bound kernels with many independent loads)
inline unsigned xcd_blocks1(long long items, int block) {
    long long b = (items + block - 1) / block;
    b = (b + XCD_COUNT - 1) / XCD_COUNT * XCD_COUNT;
    return (unsigned)b;
}
// Striped grid-wide sums: one scalar is NSTRIPE partial accumulators 128
// bytes apart, workgroup b adds into stripe b % NSTRIPE, so the per-workgroup
// atomics of a large grid spread over L2 channels instead of serialising on
// one address.  A striped scalar occupies SSIZE doubles; zero all of them.
constexpr int NSTRIPE = 32;
constexpr int SSTRIDE = 16;
constexpr int SSIZE = NSTRIPE * SSTRIDE;
__device__ inline void sadd(double* s, double v) {
    atomicAdd(s + (blockIdx.x % NSTRIPE) * SSTRIDE, v);
}
// the total, in every lane of the calling wave (all 64 lanes must call)
__device__ inline double sget(const double* s) {
    int l = threadIdx.x & 63;
    double v = l < NSTRIPE ? s[l * SSTRIDE] : 0.0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
inline double shost(const double* s) {
    double t = 0;
    for (int q = 0; q < NSTRIPE; q++) t += s[q * SSTRIDE];
    return t;
}
constexpr size_t SCRATCH_DOUBLES = 16384;  // ctx->scratch: [0, 4096) plain, [4096, ...) striped
// learned preconditioner's striped scalars in ctx->scratch: restriction sums
// {x^2, count}, line-search sums {c.res, c.A c}, then one spare (rz)
constexpr int LEARNED_WS = 4096;

// Workgroup b's logical index: with a grid that is a multiple of 8, XCD x
// (which receives workgroups x, x + 8, ...) takes the contiguous range
// [x nb / 8, (x + 1) nb / 8), so a particle kernel's field gathers and
// per-cell atomics over a cell-sorted array stay in one XCD's L2 slab
// instead of every XCD fetching the whole field.  Other grids: b.
__device__ inline size_t xcd_wg() {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    if (nb % XCD_COUNT) return b;
    return (size_t)(b % XCD_COUNT) * (nb / XCD_COUNT) + b / XCD_COUNT;
}

__device__ inline long long xcd_chunk(long long n, int block, long long& end) {
    int nb = gridDim.x;
    int b = blockIdx.x;
    int lb = (b % XCD_COUNT) * (nb / XCD_COUNT) + b / XCD_COUNT;
    long long chunk = ((n + nb - 1) / nb + block - 1) / block * block;
    long long beg = (long long)lb * chunk;
    end = beg + chunk < n ? beg + chunk : n;
    return beg;
}
