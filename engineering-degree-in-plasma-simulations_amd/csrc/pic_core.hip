// pic_core.hip -- context, geometry, species storage, transfers, diagnostics
// and profiling of the gfx950 PIC/DSMC library.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pic_internal.h"

using namespace pic;

int pic_fail(pic_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}
int pic_hip_check(pic_ctx* ctx, hipError_t e, const char* what) {
    return pic_fail(ctx, PIC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------------
// profiling: event pairs around launches, read back on demand
// ------------------------------------------------------------------------
void pic_prof_begin(pic_ctx* ctx, int cat, hipEvent_t* a, hipEvent_t* b) {
    if (!ctx->prof.on) return;
    auto take = [&]() {
        hipEvent_t e;
        if (!ctx->prof.pool.empty()) {
            e = ctx->prof.pool.back();
            ctx->prof.pool.pop_back();
        } else {
            (void)hipEventCreate(&e);
        }
        return e;
    };
    *a = take();
    *b = take();
    (void)hipEventRecord(*a, ctx->stream);
    (void)cat;
}
void pic_prof_end(pic_ctx* ctx, int cat, hipEvent_t a, hipEvent_t b) {
    if (!ctx->prof.on || !a) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->prof.recs.push_back({cat, a, b});
    if (ctx->prof.recs.size() > 8192) {  // fold finished records to bound memory
        (void)hipEventSynchronize(b);
        for (auto& r : ctx->prof.recs) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            ctx->prof.ms[r.cat] += ms;
            ctx->prof.n[r.cat] += 1;
            ctx->prof.pool.push_back(r.a);
            ctx->prof.pool.push_back(r.b);
        }
        ctx->prof.recs.clear();
    }
}

extern "C" int pic_profile_enable(pic_ctx* ctx, int on) {
    if (!ctx) return PIC_ERR_ARG;
    ctx->prof.on = on != 0;
    return PIC_OK;
}
extern "C" int pic_profile_read(pic_ctx* ctx, double ms[PIC_K_COUNT],
                                long long launches[PIC_K_COUNT]) {
    if (!ctx) return PIC_ERR_ARG;
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& r : ctx->prof.recs) {
        float m = 0;
        PIC_HIP(ctx, hipEventElapsedTime(&m, r.a, r.b));
        ctx->prof.ms[r.cat] += m;
        ctx->prof.n[r.cat] += 1;
        ctx->prof.pool.push_back(r.a);
        ctx->prof.pool.push_back(r.b);
    }
    ctx->prof.recs.clear();
    for (int c = 0; c < PIC_K_COUNT; c++) {
        if (ms) ms[c] = ctx->prof.ms[c];
        if (launches) launches[c] = ctx->prof.n[c];
    }
    return PIC_OK;
}
extern "C" int pic_profile_reset(pic_ctx* ctx) {
    if (!ctx) return PIC_ERR_ARG;
    int rc = pic_profile_read(ctx, nullptr, nullptr);
    if (rc) return rc;
    for (int c = 0; c < PIC_K_COUNT; c++) {
        ctx->prof.ms[c] = 0;
        ctx->prof.n[c] = 0;
    }
    return PIC_OK;
}

// ------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------
extern "C" const char* pic_last_error(const pic_ctx* ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int pic_create(const pic_grid_desc* gd, pic_ctx** out) {
    if (!gd || !out) return PIC_ERR_ARG;
    *out = nullptr;
    if (gd->ni < 3 || gd->nj < 3 || gd->nk < 3) return PIC_ERR_ARG;
    if (gd->ni - 1 > 2047 || gd->nj - 1 > 2047 || gd->nk - 1 > 1023) return PIC_ERR_ARG;
    // node indices are 32-bit in the grid kernels
    if ((long long)gd->ni * gd->nj * gd->nk > 0x7FFFFFFFLL) return PIC_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (!(gd->xm[d] > gd->x0[d])) return PIC_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PIC_ERR_HIP;
    pic_ctx* ctx = new pic_ctx();
    GridC& g = ctx->g;
    std::memset(&g, 0, sizeof(g));
    g.ni = gd->ni;
    g.nj = gd->nj;
    g.nk = gd->nk;
    g.ci = g.ni - 1;
    g.cj = g.nj - 1;
    g.ck = g.nk - 1;
    g.ncells = g.ci * g.cj * g.ck;
    g.nn = (long long)g.ni * g.nj * g.nk;
    int nnode[3] = {g.ni, g.nj, g.nk};
    for (int d = 0; d < 3; d++) {  // World.cpp:63-78
        g.x0[d] = gd->x0[d];
        g.dx[d] = (gd->xm[d] - gd->x0[d]) / (nnode[d] - 1);
        g.inv_dx[d] = 1 / g.dx[d];
    }
    size_t nn = (size_t)g.nn;
    ctx->object_id.assign(nn, 0);
    ctx->node_type_in.assign(nn, 0);
    ctx->phi_fixed.assign(nn, 0.0);
    hipError_t e = hipStreamCreate(&ctx->stream);
    if (e == hipSuccess) e = hipMalloc(&ctx->phi, nn * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->rho, nn * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->ef, nn * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&ctx->emax, 4 * sizeof(float));
    if (e == hipSuccess) e = hipMemsetAsync(ctx->emax, 0, 4 * sizeof(float), ctx->stream);
    if (e == hipSuccess) e = hipMalloc(&ctx->ntype, nn);
    if (e == hipSuccess) e = hipMalloc(&ctx->scratch, SCRATCH_DOUBLES * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctx->scan_tmp, ((size_t)g.ncells / 4096 + 2) * 4);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->phi, 0, nn * sizeof(double), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->rho, 0, nn * sizeof(double), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->ef, 0, nn * sizeof(float4), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        pic_destroy(ctx);
        return PIC_ERR_HIP;
    }
    *out = ctx;
    return PIC_OK;
}

void pic_unet_free(pic_ctx* ctx);

extern "C" void pic_destroy(pic_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& s : ctx->sp) {
        for (int b = 0; b < 2; b++) (void)hipFree(s.block[b]);
        (void)hipFree(s.rank);
        (void)hipFree(s.count);
        (void)hipFree(s.off);
        (void)hipFree(s.d_n);
        (void)hipFree(s.den);
        (void)hipFree(s.mom);
        (void)hipFree(s.toff);
        (void)hipFree(s.tidx);
        (void)hipFree(s.boff);
        (void)hipFree(s.toff_n);
        (void)hipFree(s.tidx_n);
    }
    for (auto& m : ctx->mcc) {
        (void)hipFree(m.sig);
        (void)hipFree(m.d_wsv);
        (void)hipFree(m.d_stats);
        (void)hipFree(m.d_part);
        (void)hipFree(m.d_spill);
    }
    pic_unet_free(ctx);
    (void)hipFree(ctx->phi);
    (void)hipFree(ctx->phi_prev);
    (void)hipFree(ctx->sor_split);
    (void)hipFree(ctx->rho);
    (void)hipFree(ctx->ef);
    (void)hipFree(ctx->emax);
    (void)hipFree(ctx->ntype);
    (void)hipFree(ctx->scratch);
    (void)hipFree(ctx->scan_tmp);
    (void)hipFree(ctx->part);
    (void)hipFree(ctx->work);
    (void)hipFree(ctx->dep_corner);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    for (auto e : ctx->solve_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& r : ctx->prof.recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto& e : ctx->prof.pool) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int pic_set_dt(pic_ctx* ctx, double dt) {
    if (!ctx || !(dt > 0)) return PIC_ERR_ARG;
    ctx->dt = dt;
    return PIC_OK;
}

extern "C" int pic_sync(pic_ctx* ctx) {
    if (!ctx) return PIC_ERR_ARG;
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

int pic_check_sid(pic_ctx* ctx, int sid) {
    if (!ctx) return PIC_ERR_ARG;
    if (sid < 0 || sid >= (int)ctx->sp.size())
        return pic_fail(ctx, PIC_ERR_ARG, "invalid species id");
    return PIC_OK;
}

int pic_ensure_pin(pic_ctx* ctx, size_t bytes) {
    if (ctx->pin_bytes >= bytes) return PIC_OK;
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    ctx->pin = nullptr;
    ctx->pin_bytes = 0;
    PIC_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->pin), bytes, hipHostMallocDefault));
    ctx->pin_bytes = bytes;
    return PIC_OK;
}

int pic_ensure_work(pic_ctx* ctx, size_t bytes) {
    if (ctx->work_bytes >= bytes) return PIC_OK;
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(ctx->work);
    ctx->work = nullptr;
    ctx->work_bytes = 0;
    PIC_HIP(ctx, hipMalloc(&ctx->work, bytes));
    ctx->work_bytes = bytes;
    return PIC_OK;
}

// ------------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------------
static void rebuild_rect_logical(pic_ctx* ctx) {
    GridC& g = ctx->g;
    g.nrect = (int)ctx->rects.size();
    for (int r = 0; r < g.nrect; r++)
        for (int d = 0; d < 3; d++) {
            const auto& R = ctx->rects[r];
            g.rlo[r][d] = (R.c[d] - R.half[d] - g.x0[d]) * g.inv_dx[d];
            g.rhi[r][d] = (R.c[d] + R.half[d] - g.x0[d]) * g.inv_dx[d];
        }
}

extern "C" int pic_add_rectangle(pic_ctx* ctx, const double centre[3], const double sides[3],
                                 double phi) {
    if (!ctx || !centre || !sides) return PIC_ERR_ARG;
    if ((int)ctx->rects.size() >= MAX_RECT)
        return pic_fail(ctx, PIC_ERR_CAPACITY, "too many rectangles");
    pic_ctx::RectH R;
    for (int d = 0; d < 3; d++) {
        R.c[d] = centre[d];
        R.half[d] = std::fabs(sides[d]) * 0.5;
    }
    R.phi = phi;
    ctx->rects.push_back(R);
    rebuild_rect_logical(ctx);
    // computeObjectID (World.cpp:276-292): nodes inside become Dirichlet at phi
    const GridC& g = ctx->g;
    std::vector<double> hphi((size_t)g.nn);
    PIC_HIP(ctx, hipMemcpyAsync(hphi.data(), ctx->phi, (size_t)g.nn * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++)
            for (int i = 0; i < g.ni; i++) {
                double x[3] = {g.x0[0] + i * g.dx[0], g.x0[1] + j * g.dx[1], g.x0[2] + k * g.dx[2]};
                bool in = true;
                for (int d = 0; d < 3; d++)
                    if (std::fabs(x[d] - R.c[d]) > R.half[d]) in = false;
                if (!in) continue;
                size_t u = ((size_t)k * g.nj + j) * g.ni + i;
                ctx->object_id[u] = 1;
                ctx->node_type_in[u] = 2;
                ctx->phi_fixed[u] = phi;
                hphi[u] = phi;
            }
    PIC_HIP(ctx, hipMemcpyAsync(ctx->phi, hphi.data(), (size_t)g.nn * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->ntype_dirty = true;
    return PIC_OK;
}

extern "C" int pic_dirichlet_walls(pic_ctx* ctx) {
    if (!ctx) return PIC_ERR_ARG;
    const GridC& g = ctx->g;
    for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++)
            for (int i = 0; i < g.ni; i++)
                if (i == 0 || i == g.ni - 1 || j == 0 || j == g.nj - 1 || k == 0 || k == g.nk - 1)
                    ctx->node_type_in[((size_t)k * g.nj + j) * g.ni + i] = 2;
    ctx->ntype_dirty = true;
    return PIC_OK;
}

// operator row type per node (PotentialSolver.cpp:414-471): Dirichlet for
// objects and flagged nodes, Neumann on the remaining faces (priority
// i-, i+, j-, j+, k-, k+), regular inside.
int pic_ensure_ntype(pic_ctx* ctx) {
    if (!ctx->ntype_dirty) return PIC_OK;
    const GridC& g = ctx->g;
    std::vector<uint8_t> nt((size_t)g.nn);
    bool neumann = false;
    for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++)
            for (int i = 0; i < g.ni; i++) {
                size_t u = ((size_t)k * g.nj + j) * g.ni + i;
                uint8_t t;
                if (ctx->object_id[u] > 0 || ctx->node_type_in[u] == 2)
                    t = NT_DIR;
                else if (i == 0)
                    t = NT_NEU_IM;
                else if (i == g.ni - 1)
                    t = NT_NEU_IP;
                else if (j == 0)
                    t = NT_NEU_JM;
                else if (j == g.nj - 1)
                    t = NT_NEU_JP;
                else if (k == 0)
                    t = NT_NEU_KM;
                else if (k == g.nk - 1)
                    t = NT_NEU_KP;
                else
                    t = NT_REG;
                nt[u] = t;
                neumann |= t != NT_REG && t != NT_DIR;
            }
    PIC_HIP(ctx, hipMemcpyAsync(ctx->ntype, nt.data(), nt.size(), hipMemcpyHostToDevice,
                                ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->ntype_dirty = false;
    ctx->has_neumann = neumann;
    return PIC_OK;
}

extern "C" int pic_get_node_flags(pic_ctx* ctx, int32_t* object_id, int32_t* node_type) {
    if (!ctx) return PIC_ERR_ARG;
    const GridC& g = ctx->g;
    for (size_t u = 0; u < (size_t)g.nn; u++) {
        if (object_id) object_id[u] = ctx->object_id[u];
        if (node_type) {
            // after buildMatrix the reference stores REGULAR/NEUMANN/DIRICHLET
            int i = (int)(u % g.ni), j = (int)((u / g.ni) % g.nj), k = (int)(u / ((size_t)g.ni * g.nj));
            int t;
            if (ctx->object_id[u] > 0 || ctx->node_type_in[u] == 2)
                t = 2;
            else if (i == 0 || i == g.ni - 1 || j == 0 || j == g.nj - 1 || k == 0 || k == g.nk - 1)
                t = 1;
            else
                t = 0;
            node_type[u] = t;
        }
    }
    return PIC_OK;
}

// ------------------------------------------------------------------------
// species
// ------------------------------------------------------------------------
extern "C" int pic_species_add(pic_ctx* ctx, double mass, double charge, double mpw0,
                               size_t capacity, int* sid) {
    if (!ctx || !sid || !(mass > 0) || capacity == 0) return PIC_ERR_ARG;
    if ((int)ctx->sp.size() >= MAX_SPECIES)
        return pic_fail(ctx, PIC_ERR_CAPACITY, "too many species");
    if (capacity > 0xF0000000ull) return pic_fail(ctx, PIC_ERR_ARG, "capacity above 2^32");
    Species s;
    s.mass = mass;
    s.charge = charge;
    s.mpw0 = mpw0;
    s.cap = capacity;
    s.seq = (uint32_t)ctx->sp.size() + 1;
    const GridC& g = ctx->g;
    size_t c = capacity;
    size_t per = c * 4;  // bytes of one 32-bit array
    size_t blk = 8 * ((per + 255) / 256 * 256);
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2 && e == hipSuccess; b++) {
        e = hipMalloc(&s.block[b], blk);
        if (e != hipSuccess) break;
        char* base = (char*)s.block[b];
        size_t stride = (per + 255) / 256 * 256;
        s.buf[b].cell = (uint32_t*)(base + 0 * stride);
        for (int d = 0; d < 3; d++) {
            s.buf[b].f[d] = (float*)(base + (1 + d) * stride);
            s.buf[b].v[d] = (float*)(base + (4 + d) * stride);
        }
        s.buf[b].w = (float*)(base + 7 * stride);
    }
    // ranks; pic_push_sort keeps 64 run bases per group of 64 particles here
    if (e == hipSuccess) e = hipMalloc(&s.rank, (c + 63) / 64 * 64 * 4);
    if (e == hipSuccess) e = hipMalloc(&s.count, ((size_t)g.ncells + 2) * 4);
    if (e == hipSuccess) e = hipMalloc(&s.off, ((size_t)g.ncells + 2) * 4);
    if (e == hipSuccess) e = hipMalloc(&s.d_n, 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&s.den, (size_t)g.nn * sizeof(double));
    if (e == hipSuccess) e = hipMemsetAsync(s.d_n, 0, 4 * sizeof(uint32_t), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s.den, 0, (size_t)g.nn * sizeof(double), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s.off, 0, ((size_t)g.ncells + 2) * 4, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        for (int b = 0; b < 2; b++) (void)hipFree(s.block[b]);
        (void)hipFree(s.rank);
        (void)hipFree(s.count);
        (void)hipFree(s.off);
        (void)hipFree(s.d_n);
        (void)hipFree(s.den);
        return pic_hip_check(ctx, e, "species allocation");
    }
    s.sorted = true;  // empty set is trivially sorted
    ctx->sp.push_back(s);
    *sid = (int)ctx->sp.size() - 1;
    return PIC_OK;
}

// physical AoS (FP64) -> packed cell + in-cell offset (World::XtoL,
// World.cpp:123-127; addParticle's bounds/object rejection, Species.cpp:420-426)
__device__ inline uint32_t locate(const GridC& g, const double* x, float* f) {
    double lc[3];
    int ijk[3];
    int lim[3] = {g.ci, g.cj, g.ck};
    for (int d = 0; d < 3; d++) {
        if (!(x[d] == x[d])) return DEAD;
        lc[d] = (x[d] - g.x0[d]) * g.inv_dx[d];
        if (lc[d] < 0.0) return DEAD;
        ijk[d] = (int)lc[d];
        if (ijk[d] >= lim[d]) return DEAD;
        float fr = (float)(lc[d] - ijk[d]);
        f[d] = fr < 1.0f ? fr : 0.99999994f;
    }
    for (int r = 0; r < g.nrect; r++) {
        bool in = true;
        for (int d = 0; d < 3; d++) {
            if (lc[d] < g.rlo[r][d] || lc[d] > g.rhi[r][d]) in = false;
        }
        if (in) return DEAD;
    }
    return pack_cell(ijk[0], ijk[1], ijk[2]);
}

__global__ void k_upload(GridC g, SpeciesBuf b, const double* __restrict__ aos, size_t n,
                         size_t base) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double* q = aos + 7 * p;
    float f[3] = {0, 0, 0};
    uint32_t c = locate(g, q, f);
    size_t o = base + p;
    b.cell[o] = c;
    for (int d = 0; d < 3; d++) {
        b.f[d][o] = f[d];
        b.v[d][o] = (float)q[3 + d];
    }
    b.w[o] = (float)q[6];
}

// record p of the download: particle p, or for a species with an indexed
// appended part (tidx) and p >= bulk its (p - bulk)-th index entry (the
// order of the cells' pick pools)
__global__ void k_download(GridC g, SpeciesBuf b, double* __restrict__ aos, size_t n,
                           const uint32_t* __restrict__ tidx, size_t bulk, const uint32_t* __restrict__ tlive) {
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    size_t p = r;
    uint32_t c = DEAD;
    if (!tidx || r < bulk) {
        c = b.cell[p];
    } else if (r - bulk < *tlive) {
        p = tidx[r - bulk];
        c = b.cell[p];
    }
    double* q = aos + 7 * r;
    if (c == DEAD) {
        for (int d = 0; d < 7; d++) q[d] = __builtin_nan("");
        return;
    }
    int ijk[3];
    unpack_cell(c, ijk[0], ijk[1], ijk[2]);
    for (int d = 0; d < 3; d++) {
        q[d] = g.x0[d] + ((double)ijk[d] + (double)b.f[d][p]) * g.dx[d];  // World::LtoX
        q[3 + d] = (double)b.v[d][p];
    }
    q[6] = (double)b.w[p];
}

extern "C" int pic_particles_upload(pic_ctx* ctx, int sid, const double* aos7, size_t n) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    Species& s = ctx->sp[sid];
    if (n > s.cap) return pic_fail(ctx, PIC_ERR_CAPACITY, "upload exceeds species capacity");
    if (n && !aos7) return PIC_ERR_ARG;
    double* stage = nullptr;
    if (n) {
        PIC_HIP(ctx, hipMalloc(&stage, n * 7 * sizeof(double)));
        hipError_t e = hipMemcpyAsync(stage, aos7, n * 7 * sizeof(double), hipMemcpyHostToDevice,
                                      ctx->stream);
        if (e == hipSuccess) {
            k_upload<<<grid_for(n, 256), 256, 0, ctx->stream>>>(ctx->g, s.buf[s.cur], stage, n, 0);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        (void)hipFree(stage);
        if (e != hipSuccess) return pic_hip_check(ctx, e, "particle upload");
    }
    uint32_t hn[4] = {(uint32_t)n, 0, 0, 0};
    PIC_HIP(ctx, hipMemcpyAsync(s.d_n, hn, sizeof(hn), hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s.n_upper = n;
    s.hist_valid = false;
    s.sorted = false;
    s.two_part = false;
    s.tail_ok = false;
    s.scrambled = true;  // the caller's order
    return PIC_OK;
}

static int read_count(pic_ctx* ctx, Species& s, size_t* n) {
    uint32_t hn[4];
    PIC_HIP(ctx, hipMemcpyAsync(hn, s.d_n, sizeof(hn), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *n = hn[0];
    s.n_upper = hn[0];
    s.bulk_n = hn[3];  // the live count of the last whole sort
    return PIC_OK;
}

extern "C" int pic_particles_count(pic_ctx* ctx, int sid, size_t* n) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    if (!n) return PIC_ERR_ARG;
    Species& s = ctx->sp[sid];
    size_t total;
    rc = read_count(ctx, s, &total);
    if (rc) return rc;
    if (s.sorted || total == 0) {
        *n = total;
        return PIC_OK;
    }
    // removed particles are still in the arrays until the next sort
    std::vector<uint32_t> cells(total);
    PIC_HIP(ctx, hipMemcpyAsync(cells.data(), s.buf[s.cur].cell, total * 4, hipMemcpyDeviceToHost,
                                ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    size_t live = 0;
    for (uint32_t c : cells) live += (c != DEAD);
    *n = live;
    return PIC_OK;
}

extern "C" int pic_particles_download(pic_ctx* ctx, int sid, double* aos7, size_t capacity,
                                      size_t* n_out) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    Species& s = ctx->sp[sid];
    size_t n;
    rc = read_count(ctx, s, &n);
    if (rc) return rc;
    // DSMC births behind a sorted part: indexed first, then listed in the
    // index's order, which is the order the next DSMC step picks them in
    if (s.tail_ok && !s.sorted && n > s.bulk_n) {
        rc = pic_sort_tail(ctx, sid);
        if (rc) return rc;
    }
    const bool perm = s.tail_ok && !s.sorted && s.two_part;
    std::vector<double> host(n * 7);
    if (n) {
        double* stage = nullptr;
        PIC_HIP(ctx, hipMalloc(&stage, n * 7 * sizeof(double)));
        k_download<<<grid_for(n, 256), 256, 0, ctx->stream>>>(ctx->g, s.buf[s.cur], stage, n,
                                                              perm ? s.tidx : nullptr, s.bulk_n,
                                                              perm ? s.toff + ctx->g.ncells : nullptr);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(host.data(), stage, n * 7 * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        (void)hipFree(stage);
        if (e != hipSuccess) return pic_hip_check(ctx, e, "particle download");
    }
    size_t live = 0;
    for (size_t p = 0; p < n; p++) {
        if (std::isnan(host[7 * p])) continue;
        if (aos7 && live < capacity) std::memcpy(aos7 + 7 * live, &host[7 * p], 7 * sizeof(double));
        live++;
    }
    if (n_out) *n_out = live;
    if (aos7 && live > capacity)
        return pic_fail(ctx, PIC_ERR_CAPACITY, "download buffer too small");
    return PIC_OK;
}

// ------------------------------------------------------------------------
// field transfer
// ------------------------------------------------------------------------
__global__ void k_ef_pack(GridC g, const double* __restrict__ e3, float4* __restrict__ ef, float* __restrict__ emax) {
    long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
    if (u < g.nn) {
        e = make_float4((float)e3[3 * u], (float)e3[3 * u + 1], (float)e3[3 * u + 2], 0.f);
        ef[u] = e;
    }
    block_max3_to(fabsf(e.x), fabsf(e.y), fabsf(e.z), emax);
}

// FP64 E = -grad(phi) (PotentialSolver.cpp:354-408), for downloads/diagnostics
__device__ inline void ef64_at(const GridC& g, const double* __restrict__ phi, int i, int j, int k,
                               double e[3]) {
    long long u = ((long long)k * g.nj + j) * g.ni + i;
    int ijk[3] = {i, j, k};
    int n[3] = {g.ni, g.nj, g.nk};
    long long st[3] = {1, g.ni, (long long)g.ni * g.nj};
    for (int d = 0; d < 3; d++) {
        double inv2 = 1.0 / (2 * g.dx[d]);
        long long s = st[d];
        if (ijk[d] == 0)
            e[d] = (3 * phi[u] - 4 * phi[u + s] + phi[u + 2 * s]) * inv2;
        else if (ijk[d] == n[d] - 1)
            e[d] = (-phi[u - 2 * s] + 4 * phi[u - s] - 3 * phi[u]) * inv2;
        else
            e[d] = (phi[u - s] - phi[u + s]) * inv2;
    }
}

__global__ void k_ef64(GridC g, const double* __restrict__ phi, double* __restrict__ out) {
    long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= g.nn) return;
    int i = (int)(u % g.ni), j = (int)((u / g.ni) % g.nj), k = (int)(u / ((long long)g.ni * g.nj));
    double e[3];
    ef64_at(g, phi, i, j, k, e);
    out[3 * u] = e[0];
    out[3 * u + 1] = e[1];
    out[3 * u + 2] = e[2];
}

extern "C" int pic_field_upload(pic_ctx* ctx, int which, const double* src) {
    if (!ctx || !src) return PIC_ERR_ARG;
    size_t nn = (size_t)ctx->g.nn;
    if (which == PIC_FIELD_PHI) {
        PIC_HIP(ctx, hipMemcpyAsync(ctx->phi, src, nn * 8, hipMemcpyHostToDevice, ctx->stream));
        ctx->phi_prev_valid = false;  // the solution history no longer leads to this phi
    } else if (which == PIC_FIELD_RHO) {
        PIC_HIP(ctx, hipMemcpyAsync(ctx->rho, src, nn * 8, hipMemcpyHostToDevice, ctx->stream));
    } else if (which == PIC_FIELD_EF) {
        int rc = pic_ensure_work(ctx, nn * 24);
        if (rc) return rc;
        PIC_HIP(ctx, hipMemcpyAsync(ctx->work, src, nn * 24, hipMemcpyHostToDevice, ctx->stream));
        PIC_HIP(ctx, hipMemsetAsync(ctx->emax, 0, 4 * sizeof(float), ctx->stream));
        k_ef_pack<<<grid_for(nn, 256), 256, 0, ctx->stream>>>(ctx->g, ctx->work, ctx->ef, ctx->emax);
        PIC_HIP(ctx, hipGetLastError());
    } else {
        return pic_fail(ctx, PIC_ERR_ARG, "field not uploadable");
    }
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

extern "C" int pic_field_download(pic_ctx* ctx, int which, double* dst) {
    if (!ctx || !dst) return PIC_ERR_ARG;
    const GridC& g = ctx->g;
    size_t nn = (size_t)g.nn;
    if (which == PIC_FIELD_PHI) {
        PIC_HIP(ctx, hipMemcpyAsync(dst, ctx->phi, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
    } else if (which == PIC_FIELD_RHO) {
        PIC_HIP(ctx, hipMemcpyAsync(dst, ctx->rho, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
    } else if (which == PIC_FIELD_EF) {
        int rc = pic_ensure_work(ctx, nn * 24);
        if (rc) return rc;
        k_ef64<<<grid_for(nn, 256), 256, 0, ctx->stream>>>(g, ctx->phi, ctx->work);
        PIC_HIP(ctx, hipGetLastError());
        PIC_HIP(ctx, hipMemcpyAsync(dst, ctx->work, nn * 24, hipMemcpyDeviceToHost, ctx->stream));
    } else if (which == PIC_FIELD_NODE_VOL) {
        for (int k = 0; k < g.nk; k++)
            for (int j = 0; j < g.nj; j++)
                for (int i = 0; i < g.ni; i++) {
                    double v = g.dx[0] * g.dx[1] * g.dx[2];
                    if (i == 0 || i == g.ni - 1) v *= 0.5;
                    if (j == 0 || j == g.nj - 1) v *= 0.5;
                    if (k == 0 || k == g.nk - 1) v *= 0.5;
                    dst[((size_t)k * g.nj + j) * g.ni + i] = v;
                }
        return PIC_OK;
    } else {
        return pic_fail(ctx, PIC_ERR_ARG, "unknown field");
    }
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

extern "C" int pic_species_field_download(pic_ctx* ctx, int sid, int which, double* dst) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    if (!dst) return PIC_ERR_ARG;
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    size_t nn = (size_t)g.nn;
    if (which == PIC_SFIELD_DEN) {
        PIC_HIP(ctx, hipMemcpyAsync(dst, s.den, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
    } else if (which >= PIC_SFIELD_NSUM && which <= PIC_SFIELD_NWW) {
        if (!s.mom) {
            size_t cnt = which == PIC_SFIELD_NVSUM ? 3 * nn : nn;
            std::fill(dst, dst + cnt, 0.0);
            return PIC_OK;
        }
        if (which == PIC_SFIELD_NVSUM) {
            std::vector<double> h(3 * nn);
            PIC_HIP(ctx, hipMemcpyAsync(h.data(), s.mom + nn, 3 * nn * 8, hipMemcpyDeviceToHost,
                                        ctx->stream));
            PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
            for (size_t u = 0; u < nn; u++)
                for (int d = 0; d < 3; d++) dst[3 * u + d] = h[d * nn + u];
            return PIC_OK;
        }
        size_t slot = which == PIC_SFIELD_NSUM ? 0 : (size_t)(which - PIC_SFIELD_NUU) + 4;
        PIC_HIP(ctx, hipMemcpyAsync(dst, s.mom + slot * nn, nn * 8, hipMemcpyDeviceToHost,
                                    ctx->stream));
    } else if (which == PIC_SFIELD_MACRO_COUNT) {
        if (!s.sorted) {
            rc = pic_sort(ctx, sid);
            if (rc) return rc;
        }
        std::vector<uint32_t> off((size_t)g.ncells + 1);
        PIC_HIP(ctx, hipMemcpyAsync(off.data(), s.off, off.size() * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
        PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < g.ncells; c++) dst[c] = (double)(off[c + 1] - off[c]);
        return PIC_OK;
    } else {
        return pic_fail(ctx, PIC_ERR_ARG, "unknown species field");
    }
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

// ------------------------------------------------------------------------
// diagnostics (Species.cpp:732-763, World.cpp:108-118): FP64 reductions
// ------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double phi_gather(const GridC& g, const double* __restrict__ phi, int i, int j,
                                    int k, double fx, double fy, double fz) {
    long long u = ((long long)k * g.nj + j) * g.ni + i;
    long long sj = g.ni, sk = (long long)g.ni * g.nj;
    double a = (1 - fx) * (1 - fy), b = (1 - fx) * fy, c = fx * (1 - fy), d = fx * fy;
    return phi[u] * a * (1 - fz) + phi[u + sk] * a * fz + phi[u + sj] * b * (1 - fz) +
           phi[u + sj + sk] * b * fz + phi[u + 1] * c * (1 - fz) + phi[u + 1 + sk] * c * fz +
           phi[u + 1 + sj] * d * (1 - fz) + phi[u + 1 + sj + sk] * d * fz;
}

__global__ void __launch_bounds__(256) k_diag(GridC g, SpeciesBuf b, const uint32_t* d_n,
                                              const double* __restrict__ phi, double* acc) {
    uint32_t n = d_n[0];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        uint32_t c = b.cell[p];
        if (c == DEAD) continue;
        double w = b.w[p], vx = b.v[0][p], vy = b.v[1][p], vz = b.v[2][p];
        s[0] += w;
        s[1] += vx * w;
        s[2] += vy * w;
        s[3] += vz * w;
        s[4] += w * (vx * vx + vy * vy + vz * vz);
        int i, j, k;
        unpack_cell(c, i, j, k);
        s[5] += w * phi_gather(g, phi, i, j, k, b.f[0][p], b.f[1][p], b.f[2][p]);
    }
    __shared__ double red[4][6];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < 6; q++) {
        double t = wave_sum(s[q]);
        if (lane == 0) red[wv][q] = t;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double t = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                   red[3][threadIdx.x];
        atomicAdd(&acc[threadIdx.x], t);
    }
}

extern "C" int pic_diag_sums(pic_ctx* ctx, int sid, double out[6]) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    if (!out) return PIC_ERR_ARG;
    Species& s = ctx->sp[sid];
    PIC_HIP(ctx, hipMemsetAsync(ctx->scratch, 0, 6 * sizeof(double), ctx->stream));
    size_t n = s.n_upper ? s.n_upper : 1;
    unsigned blocks = std::min<unsigned>(grid_for(n, 256), 2048);
    k_diag<<<blocks, 256, 0, ctx->stream>>>(ctx->g, s.buf[s.cur], s.d_n, ctx->phi, ctx->scratch);
    PIC_HIP(ctx, hipGetLastError());
    double h[6];
    PIC_HIP(ctx, hipMemcpyAsync(h, ctx->scratch, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    out[0] = h[0];
    for (int d = 0; d < 3; d++) out[1 + d] = h[1 + d] * s.mass;
    out[4] = h[4] * s.mass * 0.5;
    out[5] = h[5] * s.charge;
    return PIC_OK;
}

__global__ void __launch_bounds__(256) k_field_energy(GridC g, const double* __restrict__ phi,
                                                      double* acc) {
    double s = 0;
    for (long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x; u < g.nn;
         u += (long long)gridDim.x * blockDim.x) {
        int i = (int)(u % g.ni), j = (int)((u / g.ni) % g.nj), k = (int)(u / ((long long)g.ni * g.nj));
        double e[3];
        ef64_at(g, phi, i, j, k, e);
        s += (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) / inv_node_vol(g, i, j, k);
    }
    __shared__ double red[4];
    double t = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(acc, red[0] + red[1] + red[2] + red[3]);
}

extern "C" int pic_field_energy(pic_ctx* ctx, double* out) {
    if (!ctx || !out) return PIC_ERR_ARG;
    PIC_HIP(ctx, hipMemsetAsync(ctx->scratch, 0, sizeof(double), ctx->stream));
    unsigned blocks = std::min<unsigned>(grid_for((size_t)ctx->g.nn, 256), 2048);
    k_field_energy<<<blocks, 256, 0, ctx->stream>>>(ctx->g, ctx->phi, ctx->scratch);
    PIC_HIP(ctx, hipGetLastError());
    double h;
    PIC_HIP(ctx, hipMemcpyAsync(&h, ctx->scratch, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = 0.5 * EPS0 * h;
    return PIC_OK;
}
