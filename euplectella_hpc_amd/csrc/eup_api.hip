// extern "C" entry points of libeup.so (declared in include/eup.h).
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>

#include <algorithm>
#include <climits>

#include "eup_internal.h"

namespace eup {

static thread_local std::string g_last_error;

int fail(Handle *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    g_last_error = msg;
    return code;
}

int hip_check(Handle *h, hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    fail(h, EUP_EHIP, msg);
    return EUP_EHIP;
}

static size_t storage_bytes(int storage)
{
    switch (storage) {
    case EUP_STORE_F64: return 8;
    case EUP_STORE_F32: return 4;
    case EUP_STORE_F16: return 2;
    case EUP_STORE_BF16: return 2;
    case EUP_STORE_E3M12: return 2;
    }
    return 0;
}

static int validate(const eup_config *c, std::string &why)
{
    char buf[256];
    if (c->l < 2 || c->m < 2 || c->n < 2) {
        snprintf(buf, sizeof buf, "lattice %d x %d x %d: every size must be >= 2", c->l, c->m, c->n);
        why = buf;
        return EUP_EINVAL;
    }
    const double ghosted = (double)(c->l + 2) * (c->m + 2) * (c->n + 2);
    if (ghosted >= 2147483647.0) {
        why = "lattice too large: (l+2)(m+2)(n+2) must stay below 2^31 sites";
        return EUP_EINVAL;
    }
    if (c->mode != EUP_MODE_PARITY && c->mode != EUP_MODE_PHYSICS && c->mode != EUP_MODE_GREY) {
        why = "mode must be EUP_MODE_PARITY, EUP_MODE_PHYSICS or EUP_MODE_GREY";
        return EUP_EINVAL;
    }
    if (c->mode == EUP_MODE_GREY && c->storage != EUP_STORE_F64 && c->storage != EUP_STORE_F32) {
        why = "grey mode requires EUP_STORE_F64 or EUP_STORE_F32";
        return EUP_EINVAL;
    }
    if (c->mode == EUP_MODE_GREY && c->fgrad != 0.0) {
        // the grey step and its adjoint have no body-force term (CYLINDER_INF runs with fgrad = 0,
        // hencol.F90:37 after build.bcond.F90:319-339 zeroes u0)
        why = "grey mode requires fgrad = 0 (no body force in the grey step and its adjoint)";
        return EUP_EINVAL;
    }
    if (storage_bytes(c->storage) == 0) {
        why = "unknown storage";
        return EUP_EINVAL;
    }
    if (c->up != EUP_FACE_PERIODIC || c->down != EUP_FACE_PERIODIC || c->left != EUP_FACE_PERIODIC ||
        c->right != EUP_FACE_PERIODIC) {
        why = "y and z faces must be periodic (CYLINDER_INF face flags, build.bcond.F90:319-339)";
        return EUP_EINVAL;
    }
    const bool open = c->front == EUP_FACE_OUTFLOW && c->rear == EUP_FACE_INFLOW;
    const bool per = c->front == EUP_FACE_PERIODIC && c->rear == EUP_FACE_PERIODIC;
    if (c->mode == EUP_MODE_PARITY) {
        if (c->storage != EUP_STORE_F64) {
            why = "parity mode requires EUP_STORE_F64";
            return EUP_EINVAL;
        }
        if (!open) {
            why = "parity mode requires front = outflow (5) and rear = inflow (4)";
            return EUP_EINVAL;
        }
    } else if (!open && !per) {
        why = "x faces must be outflow/inflow (5/4) or both periodic (0/0)";
        return EUP_EINVAL;
    }
    if (!(c->omega > 0.0 && c->omega < 2.0)) {
        why = "omega must lie in (0, 2)";
        return EUP_EINVAL;
    }
    return 0;
}

} // namespace eup

using namespace eup;

struct eup_handle : Handle {};

static void release(Handle *h)
{
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    void *ptrs[] = {h->alloc[0], h->alloc[1], h->staging, h->obs_stage, h->solid, h->solid_list, h->bb_tmp,
                    h->tile_code, h->link_mask, h->mixed_list, h->outlet[0], h->outlet[1], h->diag,
                    h->site19, h->grey};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->stream) hipStreamDestroy(h->stream);
}

namespace eup {
namespace {

// sites of one chunk of the obstacle list marked solid (indices validated on the host)
__global__ void k_mark_solid(const int32_t *__restrict__ ijk, int64_t count, int l, int m,
                             uint8_t *__restrict__ solid)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t i = ijk[3 * t] - 1, j = ijk[3 * t + 1] - 1, k = ijk[3 * t + 2] - 1;
    solid[i + (int64_t)l * (j + (int64_t)m * k)] = 1;
}

constexpr int kStatBlocks = 1024;

// number of solid sites and their bounding box (1-based, min/max per axis), per block
__global__ __launch_bounds__(256) void k_mask_stats(const uint8_t *__restrict__ solid, int l, int m,
                                                    int64_t sites, int64_t *__restrict__ part)
{
    __shared__ int64_t sc[256];
    __shared__ int sb[6][256];
    int64_t c = 0;
    int b[6] = {l + 1, 0, m + 1, 0, INT32_MAX, 0};
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < sites; s += (int64_t)gridDim.x * blockDim.x) {
        if (!solid[s]) continue;
        const int i = 1 + (int)(s % l);
        const int64_t r = s / l;
        const int j = 1 + (int)(r % m), k = 1 + (int)(r / m);
        ++c;
        b[0] = min(b[0], i); b[1] = max(b[1], i);
        b[2] = min(b[2], j); b[3] = max(b[3], j);
        b[4] = min(b[4], k); b[5] = max(b[5], k);
    }
    sc[threadIdx.x] = c;
    for (int a = 0; a < 6; ++a) sb[a][threadIdx.x] = b[a];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sc[threadIdx.x] += sc[threadIdx.x + w];
            for (int a = 0; a < 6; a += 2) {
                sb[a][threadIdx.x] = min(sb[a][threadIdx.x], sb[a][threadIdx.x + w]);
                sb[a + 1][threadIdx.x] = max(sb[a + 1][threadIdx.x], sb[a + 1][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[8 * blockIdx.x] = sc[0];
        for (int a = 0; a < 6; ++a) part[8 * blockIdx.x + 1 + a] = sb[a][0];
    }
}

} // namespace
} // namespace eup

extern "C" {

int32_t eup_abi_version(void) { return EUP_ABI_VERSION; }

const char *eup_last_error(const eup_handle *h)
{
    if (h) return h->err.c_str();
    return g_last_error.c_str();
}

int eup_create(const eup_config *cfg, eup_handle **out)
{
    if (!cfg || !out) return fail(nullptr, EUP_EINVAL, "eup_create: null argument");
    *out = nullptr;
    std::string why;
    if (validate(cfg, why)) return fail(nullptr, EUP_EINVAL, "eup_create: " + why);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return fail(nullptr, EUP_EHIP, "eup_create: no HIP device available");
    if (cfg->device < 0 || cfg->device >= ndev)
        return fail(nullptr, EUP_EINVAL, "eup_create: device ordinal out of range");
    eup_handle *h = new (std::nothrow) eup_handle();
    if (!h) return fail(nullptr, EUP_EHIP, "eup_create: out of host memory");
    h->cfg = *cfg;
    Geometry &g = h->g;
    g.l = cfg->l;
    g.m = cfg->m;
    g.n = cfg->n;
    g.sites = (int64_t)g.l * g.m * g.n;
    g.ghosted = (int64_t)(g.l + 2) * (g.m + 2) * (g.n + 2);
    g.tiles_per_row = (g.l + kTile - 1) / kTile;
    g.tiles = (int64_t)g.tiles_per_row * g.m * g.n;
    h->elem = storage_bytes(cfg->storage);
    // parity mode: the reference's ghosted arrays with the row pitch padded to 16 elements and an
    // offset of 15, so that every interior row starts on a 128-byte line (the population stride is a
    // multiple of 256 elements); other modes: dense l*m*n per population
    if (cfg->mode == EUP_MODE_PARITY) {
        h->ppitch = (g.l + 2 + 15) / 16 * 16;
        h->poff = 15;
    }
    const int64_t per = cfg->mode == EUP_MODE_PARITY
                            ? h->poff + (int64_t)h->ppitch * (g.m + 2) * (g.n + 2)
                            : g.sites;
    if (cfg->mode == EUP_MODE_PARITY && per >= 2147483647LL) {
        delete h;
        return fail(nullptr, EUP_EINVAL, "eup_create: parity-mode lattice too large (padded ghosted size >= 2^31)");
    }
    h->stride = (per + 255) / 256 * 256;
    h->tiled = cfg->mode == EUP_MODE_PHYSICS;
    int rc = 0;
    if (hip_check(h, hipSetDevice(cfg->device), "hipSetDevice") ||
        hip_check(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate") ||
        hip_check(h, hipEventCreate(&h->ev0), "hipEventCreate") ||
        hip_check(h, hipEventCreate(&h->ev1), "hipEventCreate"))
        rc = EUP_EHIP;
    // physics mode: tile-major with one guard tile on either side; other modes: SoA
    const size_t state = h->tiled ? h->elem * (size_t)kTileElems * (size_t)(g.tiles + 2)
                                  : h->elem * (size_t)h->stride * 19;
    const size_t lead = h->tiled ? h->elem * (size_t)kTileElems : 0;
    for (int b = 0; b < 2 && !rc; ++b) {
        if (hip_check(h, hipMalloc(&h->alloc[b], state), "hipMalloc populations")) rc = EUP_EHIP;
        else h->buf[b] = static_cast<char *>(h->alloc[b]) + lead;
    }
    // defined contents in the guard tiles and the padding lanes of partial tiles
    for (int b = 0; b < 2 && !rc && h->tiled; ++b)
        if (hip_check(h, hipMemset(h->alloc[b], 0, state), "hipMemset populations")) rc = EUP_EHIP;
    if (!rc && (hip_check(h, hipMalloc((void **)&h->staging, sizeof(double) * g.ghosted), "hipMalloc staging") ||
                hip_check(h, hipMalloc((void **)&h->solid, g.sites), "hipMalloc solid")))
        rc = EUP_EHIP;
    if (!rc && cfg->mode == EUP_MODE_PHYSICS) {
        const size_t rb = cfg->storage == EUP_STORE_F64 ? 8 : 4;
        for (int b = 0; b < 2 && !rc; ++b)
            if (hip_check(h, hipMalloc(&h->outlet[b], rb * 5 * (size_t)g.m * g.n), "hipMalloc outlet")) rc = EUP_EHIP;
    }
    // parity mode: the outlet equilibria of the x = l sites, written by the stream-collide for the
    // next step's ghost fill (FP64, 5 x m x n)
    if (!rc && cfg->mode == EUP_MODE_PARITY &&
        hip_check(h, hipMalloc(&h->outlet[0], sizeof(double) * 5 * (size_t)g.m * g.n), "hipMalloc outlet"))
        rc = EUP_EHIP;
    if (!rc && cfg->mode == EUP_MODE_GREY &&
        hip_check(h, hipMemset(h->solid, 0, g.sites), "hipMemset solid"))
        rc = EUP_EHIP;
    if (rc) {
        g_last_error = "eup_create: " + h->err;
        release(h);
        delete h;
        return rc;
    }
    const size_t outlet = cfg->mode == EUP_MODE_PHYSICS
                              ? 2 * (cfg->storage == EUP_STORE_F64 ? 8 : 4) * 5 * (size_t)g.m * g.n
                              : (cfg->mode == EUP_MODE_PARITY ? sizeof(double) * 5 * (size_t)g.m * g.n : 0);
    h->device_bytes = (int64_t)(2 * state + sizeof(double) * g.ghosted + g.sites + outlet);
    *out = h;
    return 0;
}

void eup_destroy(eup_handle *h)
{
    if (!h) return;
    hipSetDevice(h->cfg.device);
    release(h);
    delete h;
}

int eup_set_mask(eup_handle *h, const int32_t *obs)
{
    if (!h || !obs) return fail(h, EUP_EINVAL, "eup_set_mask: null argument");
    if (h->cfg.mode == EUP_MODE_GREY)
        return fail(h, EUP_EINVAL, "eup_set_mask: grey mode takes a solid fraction (eup_set_grey)");
    if (hip_check(h, hipSetDevice(h->cfg.device), "hipSetDevice")) return EUP_EHIP;
    const Geometry &g = h->g;
    // validate values and build the dense solid flags on the host (setup-time work)
    uint8_t *sol = nullptr;
    if (hip_check(h, hipHostMalloc((void **)&sol, g.sites), "hipHostMalloc")) return EUP_EHIP;
    int64_t ns = 0;
    int bb[6] = {g.l, 0, g.m, 0, g.n, 0};
    for (int k = 1; k <= g.n; ++k)
        for (int j = 1; j <= g.m; ++j)
            for (int i = 1; i <= g.l; ++i) {
                const int64_t o = (int64_t)(i - 1) + (int64_t)g.l * ((int64_t)(j - 1) + (int64_t)g.m * (k - 1));
                const int32_t v = obs[o];
                if (v != 0 && v != 1) {
                    hipHostFree(sol);
                    return fail(h, EUP_EINVAL, "eup_set_mask: obs values must be 0 or 1");
                }
                sol[o] = (uint8_t)v;
                if (v == 1) {
                    ++ns;
                    if (i < bb[0]) bb[0] = i;
                    if (i > bb[1]) bb[1] = i;
                    if (j < bb[2]) bb[2] = j;
                    if (j > bb[3]) bb[3] = j;
                    if (k < bb[4]) bb[4] = k;
                    if (k > bb[5]) bb[5] = k;
                }
            }
    // from here on the device-side mask tables change: a failure leaves the handle without a mask
    // (eup_step refuses it) rather than with tables that no longer match nsolid
    h->mask_set = false;
    int rc = 0;
    if (hip_check(h, hipMemcpy(h->solid, sol, g.sites, hipMemcpyHostToDevice), "copy solid"))
        rc = EUP_EHIP;
    hipHostFree(sol);
    if (rc) return rc;
    h->nsolid = ns;
    memcpy(h->bbox, bb, sizeof bb);
    rc = h->cfg.mode == EUP_MODE_PARITY ? parity_prepare_mask(h) : physics_prepare_mask(h);
    if (rc) return rc;
    h->obstacles_open = false;
    h->mask_set = true;
    return 0;
}

int eup_add_obstacles(eup_handle *h, const int32_t *ijk, int64_t count)
{
    if (!h || count < 0 || (count > 0 && !ijk)) return fail(h, EUP_EINVAL, "eup_add_obstacles: bad argument");
    if (h->cfg.mode == EUP_MODE_GREY)
        return fail(h, EUP_EINVAL, "eup_add_obstacles: grey mode takes a solid fraction (eup_set_grey)");
    if (hip_check(h, hipSetDevice(h->cfg.device), "hipSetDevice")) return EUP_EHIP;
    const Geometry &g = h->g;
    // the whole list is checked before anything changes, so a rejected call leaves the handle as
    // it was
    for (int64_t t = 0; t < count; ++t) {
        const int32_t *p = ijk + 3 * t;
        if (p[0] < 1 || p[0] > g.l || p[1] < 1 || p[1] > g.m || p[2] < 1 || p[2] > g.n)
            return fail(h, EUP_ERANGE, "eup_add_obstacles: obstacle site outside the lattice");
    }
    if (!h->obstacles_open) {
        if (hip_check(h, hipMemsetAsync(h->solid, 0, g.sites, h->stream), "hipMemset solid")) return EUP_EHIP;
        h->obstacles_open = true;
        h->mask_set = false;
    }
    constexpr int64_t kChunk = 1 << 20; // triplets per staged chunk (12 MiB)
    if (!h->obs_stage &&
        hip_check(h, hipMalloc((void **)&h->obs_stage, sizeof(int32_t) * 3 * kChunk), "hipMalloc obstacle stage"))
        return EUP_EHIP;
    for (int64_t c0 = 0; c0 < count; c0 += kChunk) {
        const int64_t c = count - c0 < kChunk ? count - c0 : kChunk;
        const int32_t *p = ijk + 3 * c0;
        // the chunk is staged and marked on the handle's stream; the host buffer and the stage are
        // reused only after the stream has drained
        if (hip_check(h, hipMemcpyAsync(h->obs_stage, p, sizeof(int32_t) * 3 * c, hipMemcpyHostToDevice, h->stream),
                      "copy obstacles"))
            return EUP_EHIP;
        hipLaunchKernelGGL(k_mark_solid, dim3((unsigned)((c + 255) / 256)), dim3(256), 0, h->stream, h->obs_stage, c,
                           g.l, g.m, h->solid);
        if (hip_check(h, hipGetLastError(), "mark obstacles") || hip_check(h, hipStreamSynchronize(h->stream), "sync"))
            return EUP_EHIP;
    }
    return 0;
}

int eup_commit_obstacles(eup_handle *h)
{
    if (!h) return fail(nullptr, EUP_EINVAL, "eup_commit_obstacles: null handle");
    if (h->cfg.mode == EUP_MODE_GREY)
        return fail(h, EUP_EINVAL, "eup_commit_obstacles: grey mode takes a solid fraction (eup_set_grey)");
    if (hip_check(h, hipSetDevice(h->cfg.device), "hipSetDevice")) return EUP_EHIP;
    const Geometry &g = h->g;
    h->mask_set = false; // set again only once the mask tables are rebuilt
    if (!h->obstacles_open && hip_check(h, hipMemsetAsync(h->solid, 0, g.sites, h->stream), "hipMemset solid"))
        return EUP_EHIP;
    int64_t *part = nullptr, *part_h = nullptr;
    int rc = 0;
    if (hip_check(h, hipMalloc((void **)&part, sizeof(int64_t) * 8 * kStatBlocks), "hipMalloc") ||
        hip_check(h, hipHostMalloc((void **)&part_h, sizeof(int64_t) * 8 * kStatBlocks), "hipHostMalloc"))
        rc = EUP_EHIP;
    if (!rc) {
        hipLaunchKernelGGL(k_mask_stats, dim3(kStatBlocks), dim3(256), 0, h->stream, h->solid, g.l, g.m, g.sites,
                           part);
        if (hip_check(h, hipGetLastError(), "mask stats") ||
            hip_check(h, hipMemcpyAsync(part_h, part, sizeof(int64_t) * 8 * kStatBlocks, hipMemcpyDeviceToHost,
                                        h->stream), "copy stats") ||
            hip_check(h, hipStreamSynchronize(h->stream), "sync"))
            rc = EUP_EHIP;
    }
    if (!rc) {
        // the bounding box starts from (l, 0, m, 0, n, 0), as the reference's reader does
        int64_t ns = 0;
        int bb[6] = {g.l, 0, g.m, 0, g.n, 0};
        for (int b = 0; b < kStatBlocks; ++b) {
            const int64_t *q = part_h + 8 * b;
            if (q[0] == 0) continue;
            ns += q[0];
            bb[0] = std::min(bb[0], (int)q[1]); bb[1] = std::max(bb[1], (int)q[2]);
            bb[2] = std::min(bb[2], (int)q[3]); bb[3] = std::max(bb[3], (int)q[4]);
            bb[4] = std::min(bb[4], (int)q[5]); bb[5] = std::max(bb[5], (int)q[6]);
        }
        h->nsolid = ns;
        memcpy(h->bbox, bb, sizeof bb);
    }
    hipFree(part);
    if (part_h) hipHostFree(part_h);
    if (rc) return rc;
    rc = h->cfg.mode == EUP_MODE_PARITY ? parity_prepare_mask(h) : physics_prepare_mask(h);
    if (rc) return rc;
    h->obstacles_open = false;
    h->mask_set = true;
    return 0;
}

static int check_q(eup_handle *h, int32_t q)
{
    if (q < 1 || q > 19) return fail(h, EUP_ERANGE, "population index must be 1..19");
    return 0;
}

int eup_set_pops(eup_handle *h, int32_t q, const double *a)
{
    if (!h || !a) return fail(h, EUP_EINVAL, "eup_set_pops: null argument");
    if (check_q(h, q)) return EUP_ERANGE;
    if (hip_check(h, hipSetDevice(h->cfg.device), "hipSetDevice") ||
        hip_check(h, hipMemcpyAsync(h->staging, a, sizeof(double) * h->g.ghosted, hipMemcpyHostToDevice,
                                    h->stream), "copy pops"))
        return EUP_EHIP;
    int rc = h->cfg.mode == EUP_MODE_PARITY ? parity_set_pop(h, q) : physics_set_pop(h, q);
    if (rc) return rc;
    if (hip_check(h, hipStreamSynchronize(h->stream), "sync")) return EUP_EHIP;
    h->outlet_dirty = true;
    h->pops_set |= 1u << (q - 1);
    return 0;
}

int eup_get_pops(eup_handle *h, int32_t q, double *a)
{
    if (!h || !a) return fail(h, EUP_EINVAL, "eup_get_pops: null argument");
    if (check_q(h, q)) return EUP_ERANGE;
    if (hip_check(h, hipSetDevice(h->cfg.device), "hipSetDevice")) return EUP_EHIP;
    int rc = h->cfg.mode == EUP_MODE_PARITY ? parity_get_pop(h, q) : physics_get_pop(h, q);
    if (rc) return rc;
    if (hip_check(h, hipMemcpyAsync(a, h->staging, sizeof(double) * h->g.ghosted, hipMemcpyDeviceToHost,
                                    h->stream), "copy pops") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync"))
        return EUP_EHIP;
    return 0;
}

int eup_init_equilibrium(eup_handle *h, double ux, double uy, double uz, double noise, uint64_t seed)
{
    if (!h) return fail(nullptr, EUP_EINVAL, "eup_init_equilibrium: null handle");
    if (!(ux * ux + uy * uy + uz * uz < 0.09) || !(noise >= 0.0 && noise < 0.1))
        return fail(h, EUP_EINVAL, "eup_init_equilibrium: |u| must stay below 0.3 and noise in [0, 0.1)");
    if (hip_check(h, hipSetDevice(h->cfg.device), "hipSetDevice")) return EUP_EHIP;
    int rc = init_equilibrium(h, ux, uy, uz, noise, seed);
    if (rc) return rc;
    h->pops_set = 0x7FFFFu;
    return 0;
}

int eup_set_state(eup_handle *h, const double *a)
{
    if (!h || !a) return fail(h, EUP_EINVAL, "eup_set_state: null argument");
    for (int q = 1; q <= 19; ++q) {
        int rc = eup_set_pops(h, q, a + (int64_t)(q - 1) * h->g.ghosted);
        if (rc) return rc;
    }
    return 0;
}

int eup_get_state(eup_handle *h, double *a)
{
    if (!h || !a) return fail(h, EUP_EINVAL, "eup_get_state: null argument");
    for (int q = 1; q <= 19; ++q) {
        int rc = eup_get_pops(h, q, a + (int64_t)(q - 1) * h->g.ghosted);
        if (rc) return rc;
    }
    return 0;
}

static int ready(eup_handle *h, const char *who)
{
    if (!h) return fail(nullptr, EUP_EINVAL, std::string(who) + ": null handle");
    if (h->cfg.mode == EUP_MODE_GREY) {
        if (!h->grey_set)
            return fail(h, EUP_ESTATE, std::string(who) + ": eup_set_grey has not been called");
    } else if (!h->mask_set) {
        return fail(h, EUP_ESTATE, std::string(who) + ": no obstacle mask (eup_set_mask or eup_commit_obstacles)");
    }
    if (h->pops_set != 0x7FFFFu)
        return fail(h, EUP_ESTATE, std::string(who) + ": not all 19 populations have been set");
    if (hip_check(h, hipSetDevice(h->cfg.device), "hipSetDevice")) return EUP_EHIP;
    return 0;
}

static int do_steps(eup_handle *h, int32_t nsteps)
{
    for (int32_t t = 0; t < nsteps; ++t) {
        int rc = h->cfg.mode == EUP_MODE_PARITY ? parity_step(h)
                 : h->cfg.mode == EUP_MODE_GREY ? grey_step(h)
                                                : physics_step(h);
        if (rc) return rc;
    }
    return 0;
}

int eup_set_omega(eup_handle *h, double omega)
{
    if (!h) return fail(nullptr, EUP_EINVAL, "eup_set_omega: null handle");
    if (!(omega > 0.0 && omega < 2.0)) return fail(h, EUP_EINVAL, "eup_set_omega: omega must lie in (0, 2)");
    h->cfg.omega = omega;
    return 0;
}

int eup_step(eup_handle *h, int32_t nsteps)
{
    int rc = ready(h, "eup_step");
    if (rc) return rc;
    if (nsteps < 0) return fail(h, EUP_EINVAL, "eup_step: nsteps < 0");
    rc = do_steps(h, nsteps);
    if (rc) return rc;
    return hip_check(h, hipStreamSynchronize(h->stream), "eup_step") ? EUP_EHIP : 0;
}

int eup_step_timed(eup_handle *h, int32_t nsteps, double *gpu_ms)
{
    int rc = ready(h, "eup_step_timed");
    if (rc) return rc;
    if (nsteps < 0 || !gpu_ms) return fail(h, EUP_EINVAL, "eup_step_timed: bad argument");
    if (h->cfg.mode == EUP_MODE_PHYSICS && h->outlet_dirty) {
        rc = physics_refresh_outlet(h);
        if (rc) return rc;
    }
    if (hip_check(h, hipEventRecord(h->ev0, h->stream), "hipEventRecord")) return EUP_EHIP;
    rc = do_steps(h, nsteps);
    if (rc) return rc;
    float ms = 0.f;
    if (hip_check(h, hipEventRecord(h->ev1, h->stream), "hipEventRecord") ||
        hip_check(h, hipEventSynchronize(h->ev1), "hipEventSynchronize") ||
        hip_check(h, hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime"))
        return EUP_EHIP;
    *gpu_ms = ms;
    return 0;
}

int eup_sync(eup_handle *h)
{
    if (!h) return fail(nullptr, EUP_EINVAL, "eup_sync: null handle");
    return hip_check(h, hipStreamSynchronize(h->stream), "eup_sync") ? EUP_EHIP : 0;
}

int eup_drag_cv(eup_handle *h, const int32_t box[6], double *fluxX, double *fluxY)
{
    int rc = ready(h, "eup_drag_cv");
    if (rc) return rc;
    if (!box || !fluxX || !fluxY) return fail(h, EUP_EINVAL, "eup_drag_cv: null argument");
    return diag_drag_cv(h, box, fluxX, fluxY);
}

int eup_drag_cv_noalloc(eup_handle *h, const int32_t box[6], double *fluxX, double *fluxY)
{
    int rc = ready(h, "eup_drag_cv_noalloc");
    if (rc) return rc;
    if (!box || !fluxX || !fluxY) return fail(h, EUP_EINVAL, "eup_drag_cv_noalloc: null argument");
    return diag_drag_cv_noalloc(h, box, fluxX, fluxY);
}

int eup_drag_mem(eup_handle *h, double F[3])
{
    int rc = ready(h, "eup_drag_mem");
    if (rc) return rc;
    if (h->cfg.mode == EUP_MODE_GREY)
        return fail(h, EUP_EINVAL, "eup_drag_mem: grey mode uses eup_grey_force");
    if (!F) return fail(h, EUP_EINVAL, "eup_drag_mem: null argument");
    return diag_drag_mem(h, F);
}

int eup_probe(eup_handle *h, int32_t i, int32_t j, int32_t k, double out[4])
{
    int rc = ready(h, "eup_probe");
    if (rc) return rc;
    if (!out) return fail(h, EUP_EINVAL, "eup_probe: null argument");
    if (i < 1 || i > h->g.l || j < 1 || j > h->g.m || k < 1 || k > h->g.n)
        return fail(h, EUP_ERANGE, "eup_probe: site outside the lattice");
    double x[19];
    rc = diag_site(h, i, j, k, x);
    if (rc) return rc;
    // probe_global.F90:60-77
    const double rho = x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10]
                       + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18];
    const double xj = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]);
    const double yj = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]);
    const double zj = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]);
    out[0] = xj / rho;
    out[1] = yj / rho;
    out[2] = zj / rho;
    out[3] = rho;
    return 0;
}

int eup_vtk_velocity(eup_handle *h, int32_t plane, int32_t index, float *out)
{
    int rc = ready(h, "eup_vtk_velocity");
    if (rc) return rc;
    if (!out) return fail(h, EUP_EINVAL, "eup_vtk_velocity: null argument");
    if (plane < EUP_VTK_YZ || plane > EUP_VTK_3D) return fail(h, EUP_EINVAL, "eup_vtk_velocity: unknown plane");
    const int extent = plane == EUP_VTK_YZ ? h->g.l : (plane == EUP_VTK_XZ ? h->g.m : h->g.n);
    if (plane != EUP_VTK_3D && (index < 1 || index > extent))
        return fail(h, EUP_ERANGE, "eup_vtk_velocity: plane index outside the lattice");
    return diag_vtk_velocity(h, plane, index, out);
}

int eup_vtk_vorticity(eup_handle *h, int32_t k0, float *out)
{
    int rc = ready(h, "eup_vtk_vorticity");
    if (rc) return rc;
    if (!out) return fail(h, EUP_EINVAL, "eup_vtk_vorticity: null argument");
    if (h->g.l < 3 || h->g.m < 3) return fail(h, EUP_EINVAL, "eup_vtk_vorticity: needs l, m >= 3");
    if (k0 < 1 || k0 > h->g.n) return fail(h, EUP_ERANGE, "eup_vtk_vorticity: plane index outside the lattice");
    return diag_vtk_vorticity(h, k0, out);
}

int eup_diagno(eup_handle *h, double out[5])
{
    int rc = ready(h, "eup_diagno");
    if (rc) return rc;
    if (!out) return fail(h, EUP_EINVAL, "eup_diagno: null argument");
    return diag_diagno(h, out);
}

static int grey_ready(eup_handle *h, const char *who)
{
    if (!h) return fail(nullptr, EUP_EINVAL, std::string(who) + ": null handle");
    if (h->cfg.mode != EUP_MODE_GREY) return fail(h, EUP_EINVAL, std::string(who) + ": handle is not in grey mode");
    return hip_check(h, hipSetDevice(h->cfg.device), "hipSetDevice") ? EUP_EHIP : 0;
}

int eup_set_grey(eup_handle *h, const double *s)
{
    int rc = grey_ready(h, "eup_set_grey");
    if (rc) return rc;
    if (!s) return fail(h, EUP_EINVAL, "eup_set_grey: null argument");
    for (int64_t x = 0; x < h->g.sites; ++x)
        if (!(s[x] >= 0.0 && s[x] <= 1.0)) return fail(h, EUP_EINVAL, "eup_set_grey: s must lie in [0, 1]");
    h->grey_set = false; // set again only once the field is rewritten
    rc = grey_set_field(h, s, EUP_STORE_F64, false);
    if (rc) return rc;
    h->grey_set = true;
    h->mask_set = true;
    return 0;
}

int eup_set_grey_device(eup_handle *h, const void *s, int32_t dtype)
{
    int rc = grey_ready(h, "eup_set_grey_device");
    if (rc) return rc;
    if (!s || (dtype != EUP_STORE_F64 && dtype != EUP_STORE_F32))
        return fail(h, EUP_EINVAL, "eup_set_grey_device: null pointer or dtype not F64/F32");
    h->grey_set = false;
    rc = grey_set_field(h, s, dtype, true);
    if (rc) return rc;
    h->grey_set = true;
    h->mask_set = true;
    return 0;
}

int eup_grey_force(eup_handle *h, double F[3])
{
    int rc = grey_ready(h, "eup_grey_force");
    if (rc) return rc;
    rc = ready(h, "eup_grey_force");
    if (rc) return rc;
    if (!F) return fail(h, EUP_EINVAL, "eup_grey_force: null argument");
    return grey_force(h, F);
}

static int objective_args(eup_handle *h, int32_t t0, int32_t t1, int32_t every, double norm, double *J,
                          const char *who)
{
    int rc = grey_ready(h, who);
    if (rc) return rc;
    rc = ready(h, who);
    if (rc) return rc;
    if (!J || t0 < 1 || t1 < t0 || every < 1 || !(norm != 0.0))
        return fail(h, EUP_EINVAL, std::string(who) + ": need J != NULL, 1 <= t0 <= t1, every >= 1, norm != 0");
    return 0;
}

int eup_grey_objective(eup_handle *h, int32_t t0, int32_t t1, int32_t every, double norm, double *J,
                       double *grad)
{
    int rc = objective_args(h, t0, t1, every, norm, J, "eup_grey_objective");
    if (rc) return rc;
    return grey_objective(h, t0, t1, every, norm, J, nullptr, EUP_STORE_F64, grad);
}

int eup_grey_objective_device(eup_handle *h, int32_t t0, int32_t t1, int32_t every, double norm,
                              double *J, void *grad, int32_t dtype)
{
    int rc = objective_args(h, t0, t1, every, norm, J, "eup_grey_objective_device");
    if (rc) return rc;
    if (grad && dtype != EUP_STORE_F64 && dtype != EUP_STORE_F32)
        return fail(h, EUP_EINVAL, "eup_grey_objective_device: dtype must be F64 or F32");
    return grey_objective(h, t0, t1, every, norm, J, grad, dtype, nullptr);
}

int eup_get_info(eup_handle *h, eup_info *info)
{
    if (!h || !info) return fail(h, EUP_EINVAL, "eup_get_info: null argument");
    memset(info, 0, sizeof *info);
    info->sites = h->g.sites;
    info->nsolid = h->nsolid;
    info->tiles = h->g.tiles;
    info->tiles_fluid = h->tiles_fluid;
    info->tiles_mixed = h->tiles_mixed;
    info->tiles_solid = h->tiles_solid;
    memcpy(info->bbox, h->bbox, sizeof h->bbox);
    info->bytes_per_site = (int32_t)(2 * 19 * h->elem);
    info->device_bytes = h->device_bytes;
    return 0;
}

void *eup_stream(eup_handle *h) { return h ? (void *)h->stream : nullptr; }

} // extern "C"
