// Internal state of a libeup handle and the launchers shared between the translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/eup.h"
#include "eup_lattice.h"

namespace eup {

constexpr int kTile = 64;     // sites per tile = one wavefront along x
constexpr int kBlock = 256;   // threads per block in the site kernels (4 tiles)
constexpr int kTileElems = 19 * kTile; // population values per tile in the physics layout

// tile codes of the physics layout
constexpr int32_t kTileFluid = -1; // no solid site in the tile or adjacent to it
constexpr int32_t kTileSolid = -2; // every site of the tile is solid: skipped

struct Geometry {
    int l = 0, m = 0, n = 0;
    int64_t sites = 0;            // l*m*n
    int64_t ghosted = 0;          // (l+2)(m+2)(n+2)
    int tiles_per_row = 0;        // ceil(l / 64)
    int64_t tiles = 0;            // tiles_per_row * m * n
};

struct Handle {
    eup_config cfg{};
    Geometry g;
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t elem = 8;              // bytes per stored population value
    int64_t stride = 0;           // elements between populations in an SoA state buffer
    int ppitch = 0;               // parity mode: row pitch of the ghosted layout (elements)
    int64_t poff = 0;             // parity mode: element offset of site (0, 0, 0) in a population
    bool tiled = false;           // physics mode: tile-major layout (PopIndex)
    void *alloc[2] = {nullptr, nullptr}; // device allocations of the two state buffers
    void *buf[2] = {nullptr, nullptr}; // ping-pong population buffers (past the leading guard)
    int cur = 0;                  // buf[cur] holds the current state
    double *staging = nullptr;    // ghosted FP64 staging array (one population)
    bool mask_set = false;
    uint32_t pops_set = 0;        // bit q-1 set once population q has been provided
    // mask derived data
    uint8_t *solid = nullptr;     // dense solid flags (l*m*n)
    bool obstacles_open = false;  // eup_add_obstacles called since the last commit
    int32_t *obs_stage = nullptr; // device staging of one chunk of obstacle triplets
    int64_t nsolid = 0;
    int32_t bbox[6] = {0, 0, 0, 0, 0, 0};
    // parity mode
    int32_t *solid_list = nullptr; // ghosted linear offsets of solid sites
    double *bb_tmp = nullptr;      // nsolid x 18 bounce-back values
    // physics mode
    int32_t *tile_code = nullptr;  // per tile: kTileFluid, kTileSolid or index into link_mask
    uint32_t *link_mask = nullptr; // per mixed tile: 64 link masks
    int64_t tiles_fluid = 0, tiles_mixed = 0, tiles_solid = 0;
    int32_t *mixed_list = nullptr; // tile indices of mixed tiles
    void *outlet[2] = {nullptr, nullptr}; // outlet ghost plane, 5 x m x n (compute type)
    bool outlet_dirty = true;      // the host wrote populations: physics mode recomputes its outlet
                                   // plane, parity mode refills the constant inflow plane
    size_t fused_lds = (size_t)-1; // dynamic LDS per block of the fused kernel (occupancy cap)
    int fused_remap = -1;          // fused kernel block order: 0 dispatch, 1 XCD-contiguous, 2 XCD-owned rows
    size_t parity_lds = (size_t)-1; // the same for the parity stream-collide kernel
    // diagnostics scratch
    double *diag = nullptr;       // device scratch for reductions
    int64_t diag_len = 0;
    double *site19 = nullptr;     // one site's 19 populations in FP64
    // grey mode
    void *grey = nullptr;         // solid fraction (compute type), l*m*n
    bool grey_set = false;
    int64_t device_bytes = 0;
};

int fail(Handle *h, int code, const std::string &msg);
size_t occupancy_lds_bytes(const Handle *h, int cap);
int hip_check(Handle *h, hipError_t e, const char *what);

// parity mode (eup_parity.hip)
int parity_prepare_mask(Handle *h);
int parity_step(Handle *h);

// physics mode (eup_physics.hip)
int physics_prepare_mask(Handle *h);
int physics_step(Handle *h);
int physics_refresh_outlet(Handle *h);
int physics_set_pop(Handle *h, int q); // staging (ghosted f64) -> both buffers
int physics_get_pop(Handle *h, int q); // current buffer -> staging
int parity_set_pop(Handle *h, int q);
int init_equilibrium(Handle *h, double ux, double uy, double uz, double noise, uint64_t seed);
int parity_get_pop(Handle *h, int q);

// grey mode (eup_grey.hip)
int grey_step(Handle *h);
int grey_set_field(Handle *h, const void *s, int dtype, bool on_device);
int grey_force(Handle *h, double F[3]);
int grey_objective(Handle *h, int t0, int t1, int K, double norm, double *J, void *grad_dev, int gdtype,
                   double *grad_host);

// diagnostics (eup_diag.hip)
int diag_site(Handle *h, int i, int j, int k, double out[19]);
int diag_drag_cv(Handle *h, const int32_t box[6], double *fx, double *fy);
int diag_drag_cv_noalloc(Handle *h, const int32_t box[6], double *fx, double *fy);
int diag_drag_mem(Handle *h, double F[3]);
int diag_diagno(Handle *h, double out[5]);
int diag_vtk_velocity(Handle *h, int plane, int index, float *out);
int diag_vtk_vorticity(Handle *h, int k0, float *out);

// Element index of population q of site (i, j, k) (0-based) in a dense state buffer.
//  * SoA: q * stride + off + i + pitch * (j + m * k); grey mode: pitch = l, off = 0; parity
//    mode: ghosted coordinates (0..l+1, 0..m+1, 0..n+1) with m = m + 2, pitch = the padded row
//    pitch and off the offset that puts interior site i = 1 on a 128-byte line.
//  * tile-major (physics mode): the lattice is cut into tiles of 64 consecutive sites along x
//    (tpr = ceil(l / 64) per row, the last one padded); tile t = i / 64 + tpr * (j + m * k) holds
//    its populations as 19 runs of 64 values, so population q of the site sits at
//    (t * 19 + q) * 64 + i % 64.  One guard tile precedes tile 0 and one follows the last, so that
//    a pull shifted by one site in x from the first or last tile stays inside the allocation.
struct PopIndex {
    int64_t stride, off;
    int pitch, m, tpr, tiled;
    __host__ __device__ __forceinline__ int64_t operator()(int q, int i, int j, int k) const
    {
        const int64_t row = (int64_t)j + (int64_t)m * k;
        if (tiled) {
            const int64_t t = (int64_t)(i >> 6) + (int64_t)tpr * row;
            return ((t * 19 + q) << 6) + (i & 63);
        }
        return (int64_t)q * stride + off + i + (int64_t)pitch * row;
    }
};

inline PopIndex pop_index(const Handle *h)
{
    PopIndex p;
    const bool parity = h->cfg.mode == EUP_MODE_PARITY;
    p.stride = h->stride;
    p.off = parity ? h->poff : 0;
    p.pitch = parity ? h->ppitch : h->g.l;
    p.m = parity ? h->g.m + 2 : h->g.m;
    p.tpr = h->g.tiles_per_row;
    p.tiled = h->tiled ? 1 : 0;
    return p;
}

// deviation reference of the handle: w_q, or f_eq(rho = 1, u = (u_inflow, 0, 0)) with
// EUP_FLAG_SHIFTED_DEVIATION
inline DevRef dev_ref(const Handle *h)
{
    DevRef d;
    const bool shifted = (h->cfg.flags & EUP_FLAG_SHIFTED_DEVIATION) != 0;
    const double u = shifted ? h->cfg.u_inflow : 0.0;
    const int cx[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
    for (int q = 0; q < 19; ++q) {
        const double cu = cx[q] * u;
        d.r[q] = (float)(weight<double>(q) * (1.0 + (3.0 * cu + 1.5 * (3.0 * cu * cu - u * u))));
    }
    return d;
}

// parity layout: element of ghosted site (i, j, k) in a population, relative to the offset base
inline int64_t parity_index(const Handle *h, int i, int j, int k)
{
    return (int64_t)i + (int64_t)h->ppitch * ((int64_t)j + (int64_t)(h->g.m + 2) * (int64_t)k);
}

} // namespace eup
