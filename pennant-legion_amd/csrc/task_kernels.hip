// Task entry points: one gfx950 kernel (or kernel pair) per Legion task ID on
// the PENNANT hydro cycle.  Each is the HIP counterpart of the reference's
// CUDA variant for that task ID, restated from the serial CPU task body it must
// agree with bit for bit (file:line on each kernel), laid out for CDNA:
//
//   * Zone tiles.  A task that couples a side to its zone or to the
//     neighbouring sides of its zone runs one block per tile of whole zones
//     (at most PN_TB zones and R * PN_TB sides, R = 1, 2 or 4 per task): each
//     thread takes R sides PN_TB apart and issues the loads of all of them
//     before it uses any, the side arrays are read and written fully
//     coalesced, point values are gathered once into LDS, and a side finds
//     the previous / next side of its zone and the zone's values there (the
//     reference reads mapss3/mapss4 and gathers zone values per side).
//     Zone sums, minima and maxima are
//     formed by one lane per zone reading the staged side values in
//     ascending side order from the same starting value as the CPU body, so
//     they equal the serial result bit for bit (the reference's GPU variants
//     use double atomics, src/Mesh.hh:158-182, and are order-nondeterministic).
//   * Point sums (corner mass, corner force) gather over the point-corner CSR
//     in ascending side order instead of scattering with atomics.  A lane
//     issues all loads of a point's first four corners before using any
//     (indices clamped to the point's last corner, so the loads are
//     unconditional); a point with more corners than a lane takes is summed
//     by the whole block (contributions staged in LDS, added by one thread in
//     ascending order), which keeps the pie origin's 1024 corners off the
//     critical path of one lane.
//   * The point gathers are laid out for the 8 XCDs (xcd_block): each XCD's
//     L2 holds a contiguous share of the points, so the side values that a
//     point's corners read again for the point across the zone (mapss3) are
//     still there (HBM traffic 1.02x compulsory instead of 2.1x).  The zone
//     tiles keep the plain block order, which measured faster for them.
//   * Reductions (calcDtNew, calcDvol) use wave64 shuffles and per-block
//     partials, then a one-block fold (reduce.h).
//   * Vector (double2) fields come as {x, y, stride}: every kernel that takes
//     one is instantiated for two separate arrays (stride 1), for the
//     reference's interleaved double2 instance (stride 2, y = x + 1, 16-byte
//     accesses) and for any other stride.
#include <climits>
#include <stdint.h>

#include "common.h"
#include "reduce.h"

using namespace pn;

#define PN_TB 256          // threads of a tile block; a tile holds at most PN_TB zones
#define PN_PLIGHT 4        // corners a point-gather lane sums by itself

namespace {

// ---------------------------------------------------------------------------
// Vector-field layouts

enum { VL_STRIDED = 0, VL_SOA = 1, VL_AOS = 2 };

template <int L> struct VIO;
template <> struct VIO<VL_SOA> {
    template <class V> __device__ static d2 ld(const V& v, int64_t i) { return mk(v.x[i], v.y[i]); }
    template <class V> __device__ static double ldx(const V& v, int64_t i) { return v.x[i]; }
    __device__ static void st(const pennant_vec& v, int64_t i, d2 a) { v.x[i] = a.x; v.y[i] = a.y; }
};
template <> struct VIO<VL_AOS> {
    template <class V> __device__ static d2 ld(const V& v, int64_t i) {
        const double2 t = reinterpret_cast<const double2*>(v.x)[i];
        return mk(t.x, t.y);
    }
    template <class V> __device__ static double ldx(const V& v, int64_t i) { return v.x[2 * i]; }
    __device__ static void st(const pennant_vec& v, int64_t i, d2 a) {
        reinterpret_cast<double2*>(v.x)[i] = make_double2(a.x, a.y);
    }
};
template <> struct VIO<VL_STRIDED> {
    template <class V> __device__ static d2 ld(const V& v, int64_t i) {
        return mk(v.x[i * v.stride], v.y[i * v.stride]);
    }
    template <class V> __device__ static double ldx(const V& v, int64_t i) { return v.x[i * v.stride]; }
    __device__ static void st(const pennant_vec& v, int64_t i, d2 a) {
        v.x[i * v.stride] = a.x;
        v.y[i * v.stride] = a.y;
    }
};

// ---------------------------------------------------------------------------
// Zone tiles

struct Tile {
    int z0, nz;   // zones z0 .. z0 + nz - 1
    int s0, ns;   // their sides s0 .. s0 + ns - 1
};


// A side's place in its zone: tile-local index i of the side, zone index zi
// in the tile, the zone's first side b (tile-local), its side count n and
// the side's position k.
struct SideAt {
    int i, zi, b, n, k;
};
__device__ __forceinline__ SideAt side_at(const Tile& t, const int* zsb, const unsigned char* zof,
                                          int i) {
    SideAt c;
    c.i = i;
    c.zi = zof[i];
    c.b = zsb[c.zi] - t.s0;
    c.n = zsb[c.zi + 1] - zsb[c.zi];
    c.k = i - c.b;
    return c;
}

// Wide tiles: up to R * PN_TB sides of at most PN_TB whole zones; thread t
// takes the tile's sides t, t + PN_TB, ..., t + (R - 1) PN_TB and issues the
// loads of all of them before it uses any.  The tasks are chains of
// dependent loads (zone offsets, then maps, then point values), so their
// time is set by how many chains each wave has in flight.  R is 4 where the
// tile's LDS arrays leave four blocks per CU at that width, else 2.
template <int R>
__device__ __forceinline__ Tile wide_tile_setup(const int32_t* zsbeg, int numz, int tz, int* zsb,
                                                unsigned char* zof) {
    Tile t;
    t.z0 = blockIdx.x * tz;
    t.nz = min(tz, numz - t.z0);
    if ((int)threadIdx.x <= t.nz) zsb[threadIdx.x] = zsbeg[t.z0 + threadIdx.x];
    if (threadIdx.x == 0 && t.nz == PN_TB) zsb[PN_TB] = zsbeg[t.z0 + PN_TB];
    __syncthreads();
    t.s0 = zsb[0];
    t.ns = min(zsb[t.nz] - t.s0, R * PN_TB);
    if ((int)threadIdx.x < t.nz) {
        const int b = zsb[threadIdx.x] - t.s0;
        const int e = min(zsb[threadIdx.x + 1] - t.s0, R * PN_TB);
        for (int i = b; i < e; ++i) zof[i] = (unsigned char)threadIdx.x;
    }
    __syncthreads();
    return t;
}
template <int R>
__device__ __forceinline__ int wide_prev(const SideAt& c) {
    return min(c.b + (c.k == 0 ? c.n - 1 : c.k - 1), R * PN_TB - 1);
}
template <int R>
__device__ __forceinline__ int wide_next(const SideAt& c) {
    return min(c.b + (c.k + 1 == c.n ? 0 : c.k + 1), R * PN_TB - 1);
}
template <int R>
__device__ __forceinline__ int wide_zone_end(const Tile& t, const int* zsb, int zi) {
    return min(zsb[zi + 1] - t.s0, R * PN_TB);
}
// Shared memory and indices of a wide-tile kernel; SIDE(j) is the tile-local
// index of the thread's j-th side, SIDEC(j) the mesh index it loads from
// (the tile's last side for a thread past the end, which stores nothing).
#define PN_WIDE_SHARED(R)                                              \
    __shared__ int zsb[PN_TB + 1];                                     \
    __shared__ unsigned char zof[(R) * PN_TB];                         \
    const Tile t = wide_tile_setup<R>(m.zsbeg, m.numz, tz, zsb, zof);  \
    const int lane = threadIdx.x;                                      \
    [[maybe_unused]] const bool has_zone = lane < t.nz
#define SIDE(j) (lane + (j) * PN_TB)
#define SIDEC(j) (t.s0 + min(SIDE(j), t.ns - 1))

// ---------------------------------------------------------------------------
// Mesh tasks

// calcCtrsTask, src/Mesh.cc:797-837: ex = (px1 + px2) / 2 per side; zx =
// sum of px1 * (1/n) over the zone's sides, ascending, from 0.  Wide tile.
template <int L>
__global__ void __launch_bounds__(PN_TB) k_calc_ctrs(pennant_mesh m, int tz, pennant_cvec px,
                                                     pennant_vec ex, pennant_vec zx) {
    constexpr int R = 4;
    PN_WIDE_SHARED(R);
    __shared__ d2 x1[R * PN_TB];
    int p[R];
#pragma unroll
    for (int j = 0; j < R; ++j) p[j] = m.mapsp1[SIDEC(j)];
    d2 x[R];
#pragma unroll
    for (int j = 0; j < R; ++j) x[j] = VIO<L>::ld(px, p[j]);
#pragma unroll
    for (int j = 0; j < R; ++j)
        if (SIDE(j) < t.ns) x1[SIDE(j)] = x[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            VIO<L>::st(ex, t.s0 + i, scl(0.5, add(x[j], x1[wide_next<R>(c)])));
        }
    }
    if (has_zone) {
        const int b = zsb[lane] - t.s0, e = wide_zone_end<R>(t, zsb, lane);
        const double n = (double)(zsb[lane + 1] - zsb[lane]);
        d2 acc = mk(0., 0.);
        for (int i = b; i < e; ++i) acc = add(acc, dvs(x1[i], n));
        VIO<L>::st(zx, t.z0 + lane, acc);
    }
}

// calcVolsTask, src/Mesh.cc:928-989: side areas and volumes, zone sums; the
// bad-side count (svol <= 0) as one atomic add per block (an integer sum is
// exact in any order).
template <int L>
__global__ void __launch_bounds__(PN_TB) k_calc_vols(pennant_mesh m, int tz, pennant_cvec px,
                                                     pennant_cvec zx, double* sarea, double* svol,
                                                     double* zarea, double* zvol, int* numsbad) {
    constexpr int R = 4;
    PN_WIDE_SHARED(R);
    __shared__ d2 x1[R * PN_TB], zc[PN_TB];
    __shared__ double sa_[R * PN_TB], sv_[R * PN_TB];
    int p[R];
#pragma unroll
    for (int j = 0; j < R; ++j) p[j] = m.mapsp1[SIDEC(j)];
    if (has_zone) zc[lane] = VIO<L>::ld(zx, t.z0 + lane);
    d2 x[R];
#pragma unroll
    for (int j = 0; j < R; ++j) x[j] = VIO<L>::ld(px, p[j]);
#pragma unroll
    for (int j = 0; j < R; ++j)
        if (SIDE(j) < t.ns) x1[SIDE(j)] = x[j];
    __syncthreads();
    int count = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            const double third = 1. / 3.;
            const d2 p1 = x[j], p2 = x1[wide_next<R>(c)], z = zc[c.zi];
            const double sa = 0.5 * cross(sub(p2, p1), sub(z, p1));
            const double sv = third * sa * (p1.x + p2.x + z.x);
            sarea[t.s0 + i] = sa;
            svol[t.s0 + i] = sv;
            sa_[i] = sa;
            sv_[i] = sv;
            if (sv <= 0.) count += 1;
        }
    }
    __syncthreads();
    if (has_zone) {
        const int b = zsb[lane] - t.s0, e = wide_zone_end<R>(t, zsb, lane);
        double za = 0., zv = 0.;
        for (int i = b; i < e; ++i) {
            za += sa_[i];
            zv += sv_[i];
        }
        zarea[t.z0 + lane] = za;
        zvol[t.z0 + lane] = zv;
    }
    count = block_sum(count);
    if (threadIdx.x == 0 && count > 0) atomicAdd(numsbad, count);
}

__global__ void k_zero_int(int32_t* out) { *out = 0; }

// calcSideFracsTask, src/Mesh.cc:1085-1101: smf = sarea / zarea[z]
__global__ void __launch_bounds__(PN_TB) k_calc_side_fracs(pennant_mesh m, int tz,
                                                           const double* sarea,
                                                           const double* zarea, double* smf) {
    constexpr int R = 4;
    PN_WIDE_SHARED(R);
    __shared__ double za[PN_TB];
    if (has_zone) za[lane] = zarea[t.z0 + lane];
    double sa[R];
#pragma unroll
    for (int j = 0; j < R; ++j) sa[j] = sarea[SIDEC(j)];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            smf[t.s0 + i] = sa[j] / za[c.zi];
        }
    }
}

// calcSurfVecsTask, src/Mesh.cc:1104-1123: ssurf = rotateCCW(ex - zx[z])
template <int L>
__global__ void __launch_bounds__(PN_TB) k_calc_surf_vecs(pennant_mesh m, int tz, pennant_cvec ex,
                                                          pennant_cvec zx, pennant_vec ssurf) {
    constexpr int R = 4;
    PN_WIDE_SHARED(R);
    __shared__ d2 zc[PN_TB];
    if (has_zone) zc[lane] = VIO<L>::ld(zx, t.z0 + lane);
    d2 e[R];
#pragma unroll
    for (int j = 0; j < R; ++j) e[j] = VIO<L>::ld(ex, SIDEC(j));
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            VIO<L>::st(ssurf, t.s0 + i, rot_ccw(sub(e[j], zc[c.zi])));
        }
    }
}

// calcEdgeLenTask, src/Mesh.cc:1151-1188: elen = |px[mapsp2] - px[mapsp1]|
// per side.  Sides are independent, so there is no zone tile: every lane
// takes PN_EL sides a block apart and issues all their map loads, then all
// their point loads, before it uses any (the task is a two-step dependent
// gather, so its time is the number of such chains a wave has in flight).
#define PN_EL 4
template <int L>
__global__ void __launch_bounds__(PN_BLOCK) k_calc_edge_len(int nums, const int32_t* __restrict__ mapsp1,
                                                            const int32_t* __restrict__ mapsp2,
                                                            pennant_cvec px, double* elen) {
    const int s0 = blockIdx.x * (PN_BLOCK * PN_EL) + (int)threadIdx.x;
    int p1[PN_EL], p2[PN_EL];
#pragma unroll
    for (int j = 0; j < PN_EL; ++j) {
        const int s = min(s0 + j * PN_BLOCK, nums - 1);
        p1[j] = __builtin_nontemporal_load(mapsp1 + s);
        p2[j] = __builtin_nontemporal_load(mapsp2 + s);
    }
    d2 a[PN_EL], b[PN_EL];
#pragma unroll
    for (int j = 0; j < PN_EL; ++j) {
        a[j] = VIO<L>::ld(px, p1[j]);
        b[j] = VIO<L>::ld(px, p2[j]);
    }
#pragma unroll
    for (int j = 0; j < PN_EL; ++j) {
        const int s = s0 + j * PN_BLOCK;
        if (s < nums) __builtin_nontemporal_store(len(sub(b[j], a[j])), elen + s);
    }
}

// calcCharLenTask, src/Mesh.cc:1216-1244: zdl = min(1e99, min over the
// zone's sides of fac * sarea / elen), fac = 3 for triangles, else 4
__global__ void __launch_bounds__(PN_TB) k_calc_char_len(pennant_mesh m, int tz,
                                                         const double* sarea, const double* elen,
                                                         double* zdl) {
    constexpr int R = 2;
    PN_WIDE_SHARED(R);
    __shared__ double sdl[R * PN_TB];
    double sa[R], el[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        sa[j] = sarea[SIDEC(j)];
        el[j] = elen[SIDEC(j)];
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            const double fac = (c.n == 3 ? 3. : 4.);
            sdl[i] = fac * sa[j] / el[j];
        }
    }
    __syncthreads();
    if (has_zone) {
        const int b = zsb[lane] - t.s0, e = wide_zone_end<R>(t, zsb, lane);
        double dl = 1.e99;
        for (int i = b; i < e; ++i) dl = smin(dl, sdl[i]);
        zdl[t.z0 + lane] = dl;
    }
}

// ---------------------------------------------------------------------------
// Hydro tasks

// advPosHalfTask, src/Hydro.cc:986-1006
template <int L>
__global__ void __launch_bounds__(PN_BLOCK) k_adv_pos_half(int nump, pennant_cvec px0,
                                                           pennant_cvec pu0, pennant_vec pxp,
                                                           const double* dtp) {
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= nump) return;
    const double dth = 0.5 * (*dtp);
    VIO<L>::st(pxp, p, add(VIO<L>::ld(px0, p), scl(dth, VIO<L>::ld(pu0, p))));
}

// calcRhoTask, src/Hydro.cc:1034-1054
__global__ void __launch_bounds__(PN_BLOCK) k_calc_rho(int numz, const double* zm,
                                                       const double* zvol, double* zr) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= numz) return;
    zr[z] = zm[z] / zvol[z];
}

// The heavy points of a point-gather block: points with more than
// PN_PLIGHT corners, summed by the whole block after the light ones.
struct Heavy {
    int n;
    int p[PN_BLOCK];
};

// Point gathers staged per wave.  A wave's 64 points are neighbours, so the
// sides their corners read (side s, and the zone's previous side s3 within
// maxznump - 1 of it) lie in one or two short runs of side indices (on a
// structured mesh, the zones around a stretch of one point row: one stretch
// of each of two zone rows).  wave_runs finds them from the corners' sides
// (wave minimum and maximum, widened by a zone's reach and split at the
// middle, rounded out to whole side pairs); the corner-force gather
// (k_sum_crnr_force) loads every side of the runs once with coalesced pair
// loads into LDS and sums each point's corners from LDS, in the same order and
// arithmetic as the direct gather, which a wave whose corners spread wider (a
// stretch that wraps to the next row, small meshes) still uses.  (Staging the
// corner-mass gather's smf and zone products the same way made it slower:
// the zone values are a second dependent gather.)
#define PN_SCF_CAP 544      // staged sides per wave (two runs of 65 quads and their reach)
__device__ __forceinline__ int wave_imin(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_imax(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// The staged runs [a0, a0 + n0) and [a1, a1 + n1) (n0 = 0: none, read
// directly); even starts and lengths (a pair's second side may lie past the
// mesh's last).  Called by the whole wave.
struct Runs {
    int a0, n0, a1, n1;
};
__device__ __forceinline__ Runs wave_runs(const int (&s)[PN_PLIGHT], bool light, int reach, int nums) {
    int lo = INT_MAX, hi = -1;
    if (light) {
#pragma unroll
        for (int j = 0; j < PN_PLIGHT; ++j) {
            lo = min(lo, s[j]);
            hi = max(hi, s[j]);
        }
    }
    const int smin = max(wave_imin(lo) - reach, 0) & ~1;
    const int smax = min(wave_imax(hi) + reach, nums - 1) | 1;
    Runs r = {smin, 0, 0, 0};
    if (smax < smin) return r;
    if (smax - smin < PN_SCF_CAP) {
        r.n0 = smax - smin + 1;
        return r;
    }
    const int mid = smin + (smax - smin) / 2;
    int h0 = -1, l1 = INT_MAX;
    bool across = false;
    if (light) {
#pragma unroll
        for (int j = 0; j < PN_PLIGHT; ++j) {
            if (s[j] + reach <= mid) h0 = max(h0, s[j] + reach);
            else if (s[j] - reach > mid) l1 = min(l1, s[j] - reach);
            else across = true;
        }
    }
    h0 = wave_imax(h0) | 1;
    l1 = wave_imin(l1) & ~1;
    if (!__any(across) && h0 >= smin && l1 <= smax && h0 < l1 &&
        (h0 - smin + 1) + (smax - l1 + 1) <= PN_SCF_CAP) {
        r.n0 = h0 - smin + 1;
        r.a1 = l1;
        r.n1 = smax - l1 + 1;
    }
    return r;
}
// Slot of side x in the staged runs, or -1.
__device__ __forceinline__ int run_slot(const Runs& r, int x) {
    return (x >= r.a0 && x < r.a0 + r.n0) ? x - r.a0 : ((x >= r.a1 && x < r.a1 + r.n1) ? r.n0 + x - r.a1 : -1);
}

// Sides sd and sd + 1 of a vector field (sd + 1 clamped to the last side);
// two arrays are read with one 16-byte access each.
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));
template <int L>
__device__ __forceinline__ void ld_pair(const pennant_cvec& v, int sd, int nums, d2& a, d2& b) {
    if constexpr (L == VL_SOA) {
        if (sd + 1 < nums) {
            const f64x2u x = *reinterpret_cast<const f64x2u*>(v.x + sd);
            const f64x2u y = *reinterpret_cast<const f64x2u*>(v.y + sd);
            a = mk(x[0], y[0]);
            b = mk(x[1], y[1]);
            return;
        }
    }
    a = VIO<L>::ld(v, sd);
    b = VIO<L>::ld(v, min(sd + 1, nums - 1));
}

// calcCrnrMassTask, src/Hydro.cc:1083-1117, as a gather over the point's
// corners in ascending side order: pmaswt[p] += zr*zarea*0.5*(smf[s]+smf[s3])
__device__ __forceinline__ double crnr_mass(const double* smf, const double* zr,
                                            const double* zarea, int s, int s3, int z) {
    return zr[z] * zarea[z] * 0.5 * (smf[s] + smf[s3]);
}

__global__ void __launch_bounds__(PN_BLOCK) k_calc_crnr_mass(pennant_mesh m, const double* smf,
                                                             const double* zr,
                                                             const double* zarea,
                                                             double* pmaswt) {
    __shared__ Heavy hv;
    __shared__ double part[PN_BLOCK];
    if (threadIdx.x == 0) hv.n = 0;
    __syncthreads();
    const int p = xcd_block() * PN_BLOCK + threadIdx.x;
    if (p < m.nump) {
        const int kb = m.pcbeg[p], ke = m.pcbeg[p + 1];
        if (ke - kb > PN_PLIGHT) {
            hv.p[atomicAdd(&hv.n, 1)] = p;
        } else if (ke > kb) {
            int s[PN_PLIGHT], s3[PN_PLIGHT], z[PN_PLIGHT];
#pragma unroll
            for (int j = 0; j < PN_PLIGHT; ++j) s[j] = m.pcside[min(kb + j, ke - 1)];
#pragma unroll
            for (int j = 0; j < PN_PLIGHT; ++j) {
                s3[j] = m.mapss3[s[j]];
                z[j] = m.mapsz[s[j]];
            }
            double w[PN_PLIGHT];
#pragma unroll
            for (int j = 0; j < PN_PLIGHT; ++j) w[j] = crnr_mass(smf, zr, zarea, s[j], s3[j], z[j]);
            double acc = pmaswt[p];
#pragma unroll
            for (int j = 0; j < PN_PLIGHT; ++j)
                if (kb + j < ke) acc += w[j];
            pmaswt[p] = acc;
        }
    }
    __syncthreads();
    const int nh = hv.n;
    for (int h = 0; h < nh; ++h) {
        const int q = hv.p[h];
        const int kb = m.pcbeg[q], ke = m.pcbeg[q + 1];
        double acc = 0.;
        if (threadIdx.x == 0) acc = pmaswt[q];
        for (int k0 = kb; k0 < ke; k0 += PN_BLOCK) {
            const int k = k0 + threadIdx.x;
            if (k < ke) {
                const int s = m.pcside[k];
                part[threadIdx.x] = crnr_mass(smf, zr, zarea, s, m.mapss3[s], m.mapsz[s]);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                const int cnt = min(PN_BLOCK, ke - k0);
                for (int i = 0; i < cnt; ++i) acc += part[i];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) pmaswt[q] = acc;
    }
}

// sumCrnrForceTask, src/Hydro.cc:1159-1193, as a gather:
// pf[p] += (sfp+sfq+sft)[s] - (sfp+sfq+sft)[s3] over the point's corners
template <int L>
__device__ __forceinline__ d2 crnr_force(const pennant_cvec& sfp, const pennant_cvec& sfq,
                                         const pennant_cvec& sft, int s, int s3) {
    const d2 f = add(add(VIO<L>::ld(sfp, s), VIO<L>::ld(sfq, s)), VIO<L>::ld(sft, s));
    const d2 f3 = add(add(VIO<L>::ld(sfp, s3), VIO<L>::ld(sfq, s3)), VIO<L>::ld(sft, s3));
    return sub(f, f3);
}

// One lane per point, the corners' forces sfp + sfq + sft at s minus the same
// at s3 staged per wave (wave_runs) or read directly.
template <int L>
__global__ void __launch_bounds__(PN_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_sum_crnr_force(pennant_mesh m, pennant_cvec sfp,
                                                             pennant_cvec sfq, pennant_cvec sft,
                                                             pennant_vec pf) {
    __shared__ Heavy hv;
    __shared__ d2 part[PN_BLOCK];
    __shared__ double fx[PN_BLOCK / 64][PN_SCF_CAP], fy[PN_BLOCK / 64][PN_SCF_CAP];
    if (threadIdx.x == 0) hv.n = 0;
    __syncthreads();
    const int p = xcd_block() * PN_BLOCK + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int kb = 0, ke = 0;
    if (p < m.nump) {
        kb = m.pcbeg[p];
        ke = m.pcbeg[p + 1];
    }
    if (ke - kb > PN_PLIGHT) hv.p[atomicAdd(&hv.n, 1)] = p;
    const bool light = ke > kb && ke - kb <= PN_PLIGHT;
    const d2 pf0 = light ? VIO<L>::ld(pf, p) : mk(0., 0.);
    int s[PN_PLIGHT];
#pragma unroll
    for (int j = 0; j < PN_PLIGHT; ++j) s[j] = light ? m.pcside[min(kb + j, ke - 1)] : 0;
    const Runs rn = wave_runs(s, light, m.maxznump - 1, m.nums);
    const int a0 = rn.a0, n0 = rn.n0, a1 = rn.a1, n1 = rn.n1;
    const int n = n0 + n1;
    if (n > 0) {
        // the runs' loads in two batches (each batch's loads first, then its
        // sums into LDS); lane q takes side pair q (sides 2q, 2q + 1 of the runs)
        constexpr int NT = (PN_SCF_CAP + 127) / 128, NB = (NT + 1) / 2;
#pragma unroll
        for (int t0 = 0; t0 < NT; t0 += NB) {
            d2 fp[NB][2], fq[NB][2], ft[NB][2];
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int i = min(2 * (lane + 64 * (t0 + u)), n - 2);
                const int sd = (i < n0) ? a0 + i : a1 + (i - n0);
                ld_pair<L>(sfp, sd, m.nums, fp[u][0], fp[u][1]);
                ld_pair<L>(sfq, sd, m.nums, fq[u][0], fq[u][1]);
                ld_pair<L>(sft, sd, m.nums, ft[u][0], ft[u][1]);
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int i = 2 * (lane + 64 * (t0 + u));
                if (t0 + u < NT && i < n) {
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const d2 f = add(add(fp[u][e], fq[u][e]), ft[u][e]);
                        fx[wv][i + e] = f.x;
                        fy[wv][i + e] = f.y;
                    }
                }
            }
        }
    }
    int s3[PN_PLIGHT];
#pragma unroll
    for (int j = 0; j < PN_PLIGHT; ++j) s3[j] = light ? m.mapss3[s[j]] : 0;
    if (n > 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (light) {
            d2 acc = pf0;
#pragma unroll
            for (int j = 0; j < PN_PLIGHT; ++j)
                if (kb + j < ke) {
                    const int i = run_slot(rn, s[j]), i3 = run_slot(rn, s3[j]);
                    // (s3 outside the runs only if the mesh's maxznump understates
                    // a zone: read that corner directly)
                    acc = add(acc, (i >= 0 && i3 >= 0) ? sub(mk(fx[wv][i], fy[wv][i]), mk(fx[wv][i3], fy[wv][i3]))
                                                       : crnr_force<L>(sfp, sfq, sft, s[j], s3[j]));
                }
            VIO<L>::st(pf, p, acc);
        }
    } else if (light) {
        d2 f[PN_PLIGHT];
#pragma unroll
        for (int j = 0; j < PN_PLIGHT; ++j) f[j] = crnr_force<L>(sfp, sfq, sft, s[j], s3[j]);
        d2 acc = pf0;
#pragma unroll
        for (int j = 0; j < PN_PLIGHT; ++j)
            if (kb + j < ke) acc = add(acc, f[j]);
        VIO<L>::st(pf, p, acc);
    }
    __syncthreads();
    const int nh = hv.n;
    for (int h = 0; h < nh; ++h) {
        const int q = hv.p[h];
        const int qb = m.pcbeg[q], qe = m.pcbeg[q + 1];
        d2 acc = mk(0., 0.);
        if (threadIdx.x == 0) acc = VIO<L>::ld(pf, q);
        for (int k0 = qb; k0 < qe; k0 += PN_BLOCK) {
            const int k = k0 + threadIdx.x;
            if (k < qe) {
                const int sd = m.pcside[k];
                part[threadIdx.x] = crnr_force<L>(sfp, sfq, sft, sd, m.mapss3[sd]);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                const int cnt = min(PN_BLOCK, qe - k0);
                for (int i = 0; i < cnt; ++i) acc = add(acc, part[i]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) VIO<L>::st(pf, q, acc);
    }
}

// calcAccelTask, src/Hydro.cc:1235-1254
template <int L>
__global__ void __launch_bounds__(PN_BLOCK) k_calc_accel(int nump, pennant_cvec pf,
                                                         const double* pmaswt, pennant_vec pa) {
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= nump) return;
    const double fuzz = 1.e-99;
    VIO<L>::st(pa, p, dvs(VIO<L>::ld(pf, p), smax(pmaswt[p], fuzz)));
}

// advPosFullTask, src/Hydro.cc:1281-1306
template <int L>
__global__ void __launch_bounds__(PN_BLOCK) k_adv_pos_full(int nump, pennant_cvec px0,
                                                           pennant_cvec pu0, pennant_cvec pa,
                                                           pennant_vec px, pennant_vec pu,
                                                           const double* dtp) {
    const int p = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (p >= nump) return;
    const double dt = *dtp;
    const d2 u0 = VIO<L>::ld(pu0, p);
    const d2 u = add(u0, scl(dt, VIO<L>::ld(pa, p)));
    VIO<L>::st(pu, p, u);
    VIO<L>::st(px, p, add(VIO<L>::ld(px0, p), scl(dt * 0.5, add(u0, u))));
}

// calcWorkTask, src/Hydro.cc:1339-1406: per side
// dwork = -(dt/2) * (dot(sftot, pu0+pu at p1) * pxp.x[p1]
//                    + dot(-sftot, pu0+pu at p2) * pxp.x[p2]),
// sftot = sfp + sfq; zetot += dwork and zw = sum dwork (from 0) in
// ascending side order.
template <int L>
__global__ void __launch_bounds__(PN_TB) k_calc_work(pennant_mesh m, int tz, pennant_cvec sfp,
                                                     pennant_cvec sfq, pennant_cvec pu,
                                                     pennant_cvec pu0, pennant_cvec pxp,
                                                     double* zw, double* zetot, const double* dtp) {
    constexpr int R = 2;
    PN_WIDE_SHARED(R);
    __shared__ d2 vs[R * PN_TB];
    __shared__ double xx[R * PN_TB], dw[R * PN_TB];
    int p[R];
    d2 sp[R], sq[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        p[j] = m.mapsp1[SIDEC(j)];
        sp[j] = VIO<L>::ld(sfp, SIDEC(j));
        sq[j] = VIO<L>::ld(sfq, SIDEC(j));
    }
    const double et0 = has_zone ? zetot[t.z0 + lane] : 0.;
    d2 v[R];
    double xv[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        v[j] = add(VIO<L>::ld(pu0, p[j]), VIO<L>::ld(pu, p[j]));
        xv[j] = VIO<L>::ldx(pxp, p[j]);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            vs[i] = v[j];
            xx[i] = xv[j];
        }
    }
    __syncthreads();
    const double dth = 0.5 * (*dtp);
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            const int n2 = wide_next<R>(c);
            const d2 sftot = add(sp[j], sq[j]);
            const double sd1 = dot(sftot, v[j]);
            const double sd2 = dot(neg(sftot), vs[n2]);
            dw[i] = -dth * (sd1 * xv[j] + sd2 * xx[n2]);
        }
    }
    __syncthreads();
    if (has_zone) {
        const int b = zsb[lane] - t.s0, e = wide_zone_end<R>(t, zsb, lane);
        const int z = t.z0 + lane;
        double w = 0., et = et0;
        for (int i = b; i < e; ++i) {
            et += dw[i];
            w += dw[i];
        }
        zetot[z] = et;
        zw[z] = w;
    }
}

// calcWorkRateTask, src/Hydro.cc:1483-1510
__global__ void __launch_bounds__(PN_BLOCK) k_calc_work_rate(int numz, const double* zvol0,
                                                             const double* zvol, const double* zw,
                                                             const double* zp, double* zwrate,
                                                             const double* dtp) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= numz) return;
    const double dtinv = 1. / (*dtp);
    const double dvol = zvol[z] - zvol0[z];
    zwrate[z] = (zw[z] + zp[z] * dvol) * dtinv;
}

// calcEnergyTask, src/Hydro.cc:1545-1564
__global__ void __launch_bounds__(PN_BLOCK) k_calc_energy(int numz, const double* zetot,
                                                          const double* zm, double* ze) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= numz) return;
    const double fuzz = 1.e-99;
    ze[z] = zetot[z] / (zm[z] + fuzz);
}

// calcDtTask, src/Hydro.cc:1647-1667
__global__ void k_calc_dt(const double* dtlast, const double* dtnew, const double* dvovmax,
                          double cflv, double* dtrec) {
    double r = 1.e99;
    const double a = *dtnew;
    if (a < r) r = a;
    const double b = *dtlast * cflv / *dvovmax;
    if (b < r) r = b;
    *dtrec = r;
}

// calcGlobalDtTask, src/Driver.cc:213-245
__global__ void k_calc_global_dt(const double* dtlast, const double* dtrec, const double* time,
                                 int cycle, double dtmax, double dtinit, double dtfac,
                                 double tstop, double* dtout) {
    double dt = dtmax;
    if (cycle == 0) {
        if (dtinit < dt) dt = dtinit;
    } else {
        const double dtrecover = dtfac * (*dtlast);
        if (dtrecover < dt) dt = dtrecover;
    }
    if ((tstop - *time) < dt) dt = tstop - *time;
    if (cycle > 0) {
        if (*dtrec < dt) dt = *dtrec;
    }
    *dtout = dt;
}

// ---------------------------------------------------------------------------
// PolyGas, TTS

// PolyGas::calcStateHalfTask, src/PolyGas.cc:63-108
__global__ void __launch_bounds__(PN_BLOCK) k_calc_state_half(int numz, const double* zr,
                                                              const double* zvolp,
                                                              const double* zvol0,
                                                              const double* ze,
                                                              const double* zwrate,
                                                              const double* zm, double* zp,
                                                              double* zss, double gamma,
                                                              double ssmin, const double* dtp) {
    const int z = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (z >= numz) return;
    const double dth = 0.5 * (*dtp);
    const double gm1 = gamma - 1.;
    const double ssmin2 = smax(ssmin * ssmin, 1.e-99);
    const double r = zr[z];
    const double e = smax(ze[z], 0.);
    const double p = gm1 * r * e;
    const double pre = gm1 * e;
    const double per = gm1 * r;
    const double csqd = smax(ssmin2, pre + per * p / (r * r));
    const double ss = __builtin_sqrt(csqd);
    const double minv = 1. / zm[z];
    const double dv = (zvolp[z] - zvol0[z]) * minv;
    const double bulk = r * csqd;
    const double denom = 1. + 0.5 * per * dv;
    const double src = zwrate[z] * dth * minv;
    zp[z] = p + (per * src - r * bulk * dv) / denom;
    zss[z] = ss;
}

// PolyGas::calcForceTask, src/PolyGas.cc:164-185: sf = -zp[z] * ssurf
template <int L>
__global__ void __launch_bounds__(PN_TB) k_calc_force_pgas(pennant_mesh m, int tz,
                                                           pennant_cvec ssurf, const double* zp,
                                                           pennant_vec sf) {
    constexpr int R = 4;
    PN_WIDE_SHARED(R);
    __shared__ double zpl[PN_TB];
    if (has_zone) zpl[lane] = zp[t.z0 + lane];
    d2 ss[R];
#pragma unroll
    for (int j = 0; j < R; ++j) ss[j] = VIO<L>::ld(ssurf, SIDEC(j));
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            VIO<L>::st(sf, t.s0 + i, scl(-zpl[c.zi], ss[j]));
        }
    }
}

// TTS::calcForceTask, src/TTS.cc:55-104
template <int L>
__global__ void __launch_bounds__(PN_TB) k_calc_force_tts(pennant_mesh m, int tz,
                                                          const double* sarea, const double* smf,
                                                          pennant_cvec ssurf, const double* zarea,
                                                          const double* zr, const double* zss,
                                                          pennant_vec sf, double alfa,
                                                          double ssmin) {
    constexpr int R = 2;
    PN_WIDE_SHARED(R);
    __shared__ double za[PN_TB], zrl[PN_TB], zsl[PN_TB];
    if (has_zone) {
        za[lane] = zarea[t.z0 + lane];
        zrl[lane] = zr[t.z0 + lane];
        zsl[lane] = zss[t.z0 + lane];
    }
    double sa[R], mf[R];
    d2 ss[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        sa[j] = sarea[SIDEC(j)];
        mf[j] = smf[SIDEC(j)];
        ss[j] = VIO<L>::ld(ssurf, SIDEC(j));
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            const double vfacinv = za[c.zi] / sa[j];
            const double r = zrl[c.zi];
            const double srho = r * mf[j] * vfacinv;
            double sstmp = smax(zsl[c.zi], ssmin);
            sstmp = alfa * sstmp * sstmp;
            const double sdp = sstmp * (srho - r);
            VIO<L>::st(sf, t.s0 + i, scl(-sdp, ss[j]));
        }
    }
}

// ---------------------------------------------------------------------------
// QCS

// QCS::setCornerDivTask, src/QCS.cc:116-242.  Corner c = side s2 of its
// zone; s = ss3[c] is the zone's previous side.  Its points: p = p2(s) =
// p1(c), p1 = p1(s), p2 = p2(c) = p1(ss4[c]).  Each lane stages its side's
// first point (position and velocity), edge midpoint and edge length; the
// zone velocity zuc = sum pu0[p1] * (1/n) (ascending, from 0) is formed per
// zone; then each lane computes its corner from the staged values.
template <int L>
__global__ void __launch_bounds__(PN_TB) k_set_corner_div(pennant_mesh m, int tz, pennant_cvec ex,
                                                          const double* elen, pennant_cvec zx,
                                                          pennant_cvec px, pennant_cvec pu,
                                                          pennant_vec zuc, double* carea,
                                                          double* ccos, double* cdiv,
                                                          double* cevol, double* cdu) {
    constexpr int R = 2;
    PN_WIDE_SHARED(R);
    __shared__ d2 u1[R * PN_TB], e1[R * PN_TB], zxl[PN_TB], zul[PN_TB];
    __shared__ double el[R * PN_TB];
    int p[R];
    d2 e[R];
    double l[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        p[j] = m.mapsp1[SIDEC(j)];
        e[j] = VIO<L>::ld(ex, SIDEC(j));
        l[j] = elen[SIDEC(j)];
    }
    if (has_zone) zxl[lane] = VIO<L>::ld(zx, t.z0 + lane);
    d2 u[R], x[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        u[j] = VIO<L>::ld(pu, p[j]);
        x[j] = VIO<L>::ld(px, p[j]);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            u1[i] = u[j];
            e1[i] = e[j];
            el[i] = l[j];
        }
    }
    __syncthreads();
    if (has_zone) {
        const int b = zsb[lane] - t.s0, ez = wide_zone_end<R>(t, zsb, lane);
        const double n = (double)(zsb[lane + 1] - zsb[lane]);
        d2 acc = mk(0., 0.);
        for (int i = b; i < ez; ++i) acc = add(acc, dvs(u1[i], n));
        zul[lane] = acc;
        VIO<L>::st(zuc, t.z0 + lane, acc);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i >= t.ns) continue;
        const SideAt c = side_at(t, zsb, zof, i);
        const int ip = wide_prev<R>(c), in = wide_next<R>(c);
        const int cc = t.s0 + i;
        const d2 up0 = u[j];
        const d2 xp0 = x[j];
        const d2 up1 = scl(0.5, add(up0, u1[in]));
        const d2 xp1 = e[j];
        const d2 up2 = zul[c.zi];
        const d2 xp2 = zxl[c.zi];
        const d2 up3 = scl(0.5, add(u1[ip], up0));
        const d2 xp3 = e1[ip];
        const double cvolume = 0.5 * cross(sub(xp2, xp0), sub(xp3, xp1));
        carea[cc] = cvolume;
        const d2 v1 = sub(xp3, xp0);
        const d2 v2 = sub(xp1, xp0);
        const double de1 = el[ip];
        const double de2 = l[j];
        const double minelen = smin(de1, de2);
        ccos[cc] = ((minelen < 1.e-12) ? 0. : 4. * dot(v1, v2) / (de1 * de2));
        const double div =
            (cross(sub(up2, up0), sub(xp3, xp1)) - cross(sub(up3, up1), sub(xp2, xp0))) / (2.0 * cvolume);
        cdiv[cc] = div;
        const d2 dxx1 = scl(0.5, sub(sub(add(xp1, xp2), xp0), xp3));
        const d2 dxx2 = scl(0.5, sub(sub(add(xp2, xp3), xp0), xp1));
        const double dx1 = len(dxx1);
        const double dx2 = len(dxx2);
        const d2 duav = scl(0.25, add(add(add(up0, up1), up2), up3));
        const double test1 = fabs(dot(dxx1, duav) * dx2);
        const double test2 = fabs(dot(dxx2, duav) * dx1);
        const double num = (test1 > test2 ? dx1 : dx2);
        const double den = (test1 > test2 ? dx2 : dx1);
        const double r = num / den;
        double evol = __builtin_sqrt(4.0 * cvolume * r);
        evol = smin(evol, 2.0 * minelen);
        const double dv1 = len2(sub(sub(add(up1, up2), up0), up3));
        const double dv2 = len2(sub(sub(add(up2, up3), up0), up1));
        double du = __builtin_sqrt(smax(dv1, dv2));
        evol = (div < 0.0 ? evol : 0.);
        du = (div < 0.0 ? du : 0.);
        cevol[cc] = evol;
        cdu[cc] = du;
    }
}

// QCS::setQCnForceTask, src/QCS.cc:246-327: crmu, and the viscous edge
// forces cqe1 = crmu (pu[p] - pu[p1]) / elen[s], cqe2 = crmu (pu[p2] -
// pu[p]) / elen[c] with s = ss3[c], p = p1(c), p1 = p1(s), p2 = p1(ss4[c])
template <int L>
__global__ void __launch_bounds__(PN_TB) k_set_qcn_force(pennant_mesh m, int tz, const double* elen,
                                                         const double* cdiv, const double* cdu,
                                                         const double* cevol, const double* zrp,
                                                         const double* zss, pennant_cvec pu,
                                                         double* crmu, pennant_vec cqe1,
                                                         pennant_vec cqe2, double qgamma,
                                                         double q1, double q2) {
    constexpr int R = 2;
    PN_WIDE_SHARED(R);
    __shared__ d2 u1[R * PN_TB];
    __shared__ double el[R * PN_TB], zrl[PN_TB], zsl[PN_TB];
    int p[R];
    double l[R], du[R], ev[R], dv[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        p[j] = m.mapsp1[SIDEC(j)];
        l[j] = elen[SIDEC(j)];
        du[j] = cdu[SIDEC(j)];
        ev[j] = cevol[SIDEC(j)];
        dv[j] = cdiv[SIDEC(j)];
    }
    if (has_zone) {
        zrl[lane] = zrp[t.z0 + lane];
        zsl[lane] = zss[t.z0 + lane];
    }
    d2 u[R];
#pragma unroll
    for (int j = 0; j < R; ++j) u[j] = VIO<L>::ld(pu, p[j]);
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            u1[i] = u[j];
            el[i] = l[j];
        }
    }
    __syncthreads();
    const double gammap1 = qgamma + 1.0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i >= t.ns) continue;
        const SideAt c = side_at(t, zsb, zof, i);
        const int cc = t.s0 + i;
        const int ip = wide_prev<R>(c), in = wide_next<R>(c);
        const double ztmp2 = q2 * 0.25 * gammap1 * du[j];
        const double ztmp1 = q1 * zsl[c.zi];
        const double zkur = ztmp2 + __builtin_sqrt(ztmp2 * ztmp2 + ztmp1 * ztmp1);
        const double mu = zkur * zrl[c.zi] * ev[j];
        const double cm = ((dv[j] > 0.0) ? 0. : mu);
        crmu[cc] = cm;
        VIO<L>::st(cqe1, cc, dvs(scl(cm, sub(u[j], u1[ip])), el[ip]));
        VIO<L>::st(cqe2, cc, dvs(scl(cm, sub(u1[in], u[j])), l[j]));
    }
}

// QCS::setForceTask, src/QCS.cc:331-380: each lane clamps its corner's ccos
// and computes cw, then the side force from corner c1 = s and its zone's
// next corner c2 = ss4[s], both staged in LDS.
template <int L>
__global__ void __launch_bounds__(PN_TB) k_set_force_qcs(pennant_mesh m, int tz, const double* carea,
                                                         pennant_cvec cqe1, pennant_cvec cqe2,
                                                         const double* elen, double* ccos,
                                                         double* cw, pennant_vec sfq) {
    constexpr int R = 1;
    PN_WIDE_SHARED(R);
    __shared__ d2 q1l[R * PN_TB], q2l[R * PN_TB];
    __shared__ double cwl[R * PN_TB], ccl[R * PN_TB];
    double cc0[R], ca[R], el[R];
    d2 qa[R], qb[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        cc0[j] = ccos[SIDEC(j)];
        ca[j] = carea[SIDEC(j)];
        qa[j] = VIO<L>::ld(cqe1, SIDEC(j));
        qb[j] = VIO<L>::ld(cqe2, SIDEC(j));
        el[j] = elen[SIDEC(j)];
    }
    double w[R], cl[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const double csin2 = 1.0 - cc0[j] * cc0[j];
        w[j] = ((csin2 < 1.e-4) ? 0. : ca[j] / csin2);
        cl[j] = ((csin2 < 1.e-4) ? 0. : cc0[j]);
        const int i = SIDE(j);
        if (i < t.ns) {
            cw[t.s0 + i] = w[j];
            ccos[t.s0 + i] = cl[j];
            cwl[i] = w[j];
            ccl[i] = cl[j];
            q1l[i] = qa[j];
            q2l[i] = qb[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i >= t.ns) continue;
        const SideAt c = side_at(t, zsb, zof, i);
        const int i2 = wide_next<R>(c);
        const double cw1 = w[j], cw2 = cwl[i2];
        const double ccos1 = cl[j], ccos2 = ccl[i2];
        const d2 cqe11 = qa[j], cqe12 = q1l[i2];
        const d2 cqe21 = qb[j], cqe22 = q2l[i2];
        const d2 f = add(scl(cw1, add(cqe21, scl(ccos1, cqe11))),
                         scl(cw2, add(cqe12, scl(ccos2, cqe22))));
        VIO<L>::st(sfq, t.s0 + i, dvs(f, el[j]));
    }
}

// QCS::setVelDiffTask, src/QCS.cc:384-448: ztmp = max(0, max over the
// zone's sides of |dot(du, dx)| / elen) in ascending side order;
// zdu = q1 zss + 2 q2 ztmp
template <int L>
__global__ void __launch_bounds__(PN_TB) k_set_vel_diff(pennant_mesh m, int tz, const double* elen,
                                                        const double* zss, pennant_cvec px,
                                                        pennant_cvec pu, double* ztmp, double* zdu,
                                                        double q1, double q2) {
    constexpr int R = 2;
    PN_WIDE_SHARED(R);
    __shared__ d2 x1[R * PN_TB], u1[R * PN_TB];
    __shared__ double dux_[R * PN_TB];
    int p[R];
    double l[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        p[j] = m.mapsp1[SIDEC(j)];
        l[j] = elen[SIDEC(j)];
    }
    const double zs = has_zone ? zss[t.z0 + lane] : 0.;
    d2 x[R], u[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        x[j] = VIO<L>::ld(px, p[j]);
        u[j] = VIO<L>::ld(pu, p[j]);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            x1[i] = x[j];
            u1[i] = u[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int i = SIDE(j);
        if (i < t.ns) {
            const SideAt c = side_at(t, zsb, zof, i);
            const int i2 = wide_next<R>(c);
            const d2 dx = sub(x1[i2], x[j]);
            const d2 du = sub(u1[i2], u[j]);
            const double lenx = l[j];
            double dux = dot(du, dx);
            dux = (lenx > 0. ? fabs(dux) / lenx : 0.);
            dux_[i] = dux;
        }
    }
    __syncthreads();
    if (has_zone) {
        const int b = zsb[lane] - t.s0, e = wide_zone_end<R>(t, zsb, lane);
        double tm = 0.;
        for (int i = b; i < e; ++i) tm = smax(tm, dux_[i]);
        const int z = t.z0 + lane;
        ztmp[z] = tm;
        zdu[z] = q1 * zs + 2. * q2 * tm;
    }
}

// ---------------------------------------------------------------------------
// HydroBC::applyFixedBCTask, src/HydroBC.cc:150-183
template <int L>
__global__ void __launch_bounds__(PN_BLOCK) k_apply_fixed_bc(const int32_t* mapbp, int numb,
                                                             double vx, double vy, pennant_vec pf,
                                                             pennant_vec pu0) {
    const int b = blockIdx.x * PN_BLOCK + threadIdx.x;
    if (b >= numb) return;
    const int p = mapbp[b];
    const d2 u = mk(vx, vy);
    VIO<L>::st(pu0, p, project(VIO<L>::ld(pu0, p), u));
    VIO<L>::st(pf, p, project(VIO<L>::ld(pf, p), u));
}

// ---------------------------------------------------------------------------
// Host side: argument checks and layout choice

inline int launch_status() { return (int)hipGetLastError(); }

// Every vector argument of one call, normalised (stride 0 -> 1); the call
// runs in the layout all of them share, else through the strided path.
struct Layout {
    bool soa = true, aos = true, bad = false;
    template <class V> V take(V v) {
        if (v.stride == 0) v.stride = 1;
        if (v.stride < 0) bad = true;
        if (v.stride != 1) soa = false;
        const bool inter = v.stride == 2 && v.y == v.x + 1 &&
                           (reinterpret_cast<uintptr_t>(v.x) & 15) == 0;
        if (!inter) aos = false;
        return v;
    }
    int kind() const { return soa ? VL_SOA : (aos ? VL_AOS : VL_STRIDED); }
};

// zones per tile block for the mesh; 0 when the mesh descriptor cannot be
// tiled (zones of fewer than 3 or more than PENNANT_MAX_ZSIDES sides)
inline int tile_blocks(const pennant_mesh* m, int tz) { return (m->numz + tz - 1) / tz; }
// zones of a wide tile of r * PN_TB sides: as many whole zones as it holds,
// at most one per thread
inline int wide_tile_zones(const pennant_mesh* m, int r) {
    if (m->maxznump < 3 || m->maxznump > PENNANT_MAX_ZSIDES) return 0;
    const int tz = r * PN_TB / m->maxznump;
    return tz < PN_TB ? tz : PN_TB;
}

#define PN_LAUNCH_L(lay, kern, grid, block, st, ...)                                               \
    do {                                                                                           \
        switch (lay) {                                                                             \
            case VL_SOA: hipLaunchKernelGGL(kern<VL_SOA>, grid, block, 0, st, __VA_ARGS__); break;   \
            case VL_AOS: hipLaunchKernelGGL(kern<VL_AOS>, grid, block, 0, st, __VA_ARGS__); break;   \
            default: hipLaunchKernelGGL(kern<VL_STRIDED>, grid, block, 0, st, __VA_ARGS__); break;   \
        }                                                                                          \
    } while (0)

// Common preamble of a tile-kernel entry point: nothing to do for an empty
// mesh, EINVAL for a descriptor that cannot be tiled.
#define PN_WIDE_ENTRY(m, tz, R)                       \
    if (!(m)) return PENNANT_EINVAL;                  \
    if ((m)->numz <= 0) return launch_status();       \
    const int tz = wide_tile_zones(m, R);             \
    if (tz == 0) return PENNANT_EINVAL;               \
    const dim3 tgrid(tile_blocks(m, tz)), tblk(PN_TB)

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

size_t pennant_hip_reduce_workspace_bytes(void) {
    return (size_t)PN_RED_BLOCKS * sizeof(double);
}

int pennant_hip_calc_ctrs(const pennant_mesh* m, pennant_cvec px, pennant_vec ex, pennant_vec zx,
                          hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 4);
    Layout l;
    px = l.take(px); ex = l.take(ex); zx = l.take(zx);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_calc_ctrs, tgrid, tblk, st, *m, tz, px, ex, zx);
    return launch_status();
}

int pennant_hip_calc_vols(const pennant_mesh* m, pennant_cvec px, pennant_cvec zx, double* sarea,
                          double* svol, double* zarea, double* zvol, int32_t* numsbad, void* ws,
                          hipStream_t st) {
    (void)ws;
    if (!m || !numsbad) return PENNANT_EINVAL;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, numsbad);
    PN_WIDE_ENTRY(m, tz, 4);
    Layout l;
    px = l.take(px); zx = l.take(zx);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_calc_vols, tgrid, tblk, st, *m, tz, px, zx, sarea, svol, zarea, zvol,
                numsbad);
    return launch_status();
}

int pennant_hip_calc_side_fracs(const pennant_mesh* m, const double* sarea, const double* zarea,
                                double* smf, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 4);
    hipLaunchKernelGGL(k_calc_side_fracs, tgrid, tblk, 0, st, *m, tz, sarea, zarea, smf);
    return launch_status();
}

int pennant_hip_calc_surf_vecs(const pennant_mesh* m, pennant_cvec ex, pennant_cvec zx,
                               pennant_vec ssurf, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 4);
    Layout l;
    ex = l.take(ex); zx = l.take(zx); ssurf = l.take(ssurf);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_calc_surf_vecs, tgrid, tblk, st, *m, tz, ex, zx, ssurf);
    return launch_status();
}

int pennant_hip_calc_edge_len(const pennant_mesh* m, pennant_cvec px, double* elen,
                              hipStream_t st) {
    if (!m || !m->mapsp1 || !m->mapsp2) return PENNANT_EINVAL;
    if (m->nums <= 0) return launch_status();
    Layout l;
    px = l.take(px);
    if (l.bad) return PENNANT_EINVAL;
    const dim3 grid((unsigned)((m->nums + PN_BLOCK * PN_EL - 1) / (PN_BLOCK * PN_EL)));
    PN_LAUNCH_L(l.kind(), k_calc_edge_len, grid, dim3(PN_BLOCK), st, m->nums, m->mapsp1, m->mapsp2, px,
                elen);
    return launch_status();
}

int pennant_hip_calc_char_len(const pennant_mesh* m, const double* sarea, const double* elen,
                              double* zdl, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 2);
    hipLaunchKernelGGL(k_calc_char_len, tgrid, tblk, 0, st, *m, tz, sarea, elen, zdl);
    return launch_status();
}

int pennant_hip_adv_pos_half(const pennant_mesh* m, pennant_cvec px0, pennant_cvec pu0,
                             pennant_vec pxp, const double* dt, hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    Layout l;
    px0 = l.take(px0); pu0 = l.take(pu0); pxp = l.take(pxp);
    if (l.bad) return PENNANT_EINVAL;
    if (m->nump > 0)
        PN_LAUNCH_L(l.kind(), k_adv_pos_half, dim3(nblocks(m->nump)), dim3(PN_BLOCK), st, m->nump,
                    px0, pu0, pxp, dt);
    return launch_status();
}

int pennant_hip_calc_rho(const pennant_mesh* m, const double* zm, const double* zvol, double* zr,
                         hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_rho, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, m->numz, zm,
                           zvol, zr);
    return launch_status();
}

int pennant_hip_calc_crnr_mass(const pennant_mesh* m, const double* smf, const double* zr,
                               const double* zarea, double* pmaswt, hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    if (m->nump > 0)
        hipLaunchKernelGGL(k_calc_crnr_mass, dim3(nblocks(m->nump)), dim3(PN_BLOCK), 0, st, *m,
                           smf, zr, zarea, pmaswt);
    return launch_status();
}

int pennant_hip_sum_crnr_force(const pennant_mesh* m, pennant_cvec sfp, pennant_cvec sfq,
                               pennant_cvec sft, pennant_vec pf, hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    Layout l;
    sfp = l.take(sfp); sfq = l.take(sfq); sft = l.take(sft); pf = l.take(pf);
    if (l.bad) return PENNANT_EINVAL;
    if (m->nump > 0)
        PN_LAUNCH_L(l.kind(), k_sum_crnr_force, dim3(nblocks(m->nump)), dim3(PN_BLOCK), st, *m,
                    sfp, sfq, sft, pf);
    return launch_status();
}

int pennant_hip_calc_accel(const pennant_mesh* m, pennant_cvec pf, const double* pmaswt,
                           pennant_vec pa, hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    Layout l;
    pf = l.take(pf); pa = l.take(pa);
    if (l.bad) return PENNANT_EINVAL;
    if (m->nump > 0)
        PN_LAUNCH_L(l.kind(), k_calc_accel, dim3(nblocks(m->nump)), dim3(PN_BLOCK), st, m->nump,
                    pf, pmaswt, pa);
    return launch_status();
}

int pennant_hip_adv_pos_full(const pennant_mesh* m, pennant_cvec px0, pennant_cvec pu0,
                             pennant_cvec pa, pennant_vec px, pennant_vec pu, const double* dt,
                             hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    Layout l;
    px0 = l.take(px0); pu0 = l.take(pu0); pa = l.take(pa); px = l.take(px); pu = l.take(pu);
    if (l.bad) return PENNANT_EINVAL;
    if (m->nump > 0)
        PN_LAUNCH_L(l.kind(), k_adv_pos_full, dim3(nblocks(m->nump)), dim3(PN_BLOCK), st, m->nump,
                    px0, pu0, pa, px, pu, dt);
    return launch_status();
}

int pennant_hip_calc_work(const pennant_mesh* m, pennant_cvec sfp, pennant_cvec sfq,
                          pennant_cvec pu, pennant_cvec pu0, pennant_cvec pxp, double* zw,
                          double* zetot, const double* dt, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 2);
    Layout l;
    sfp = l.take(sfp); sfq = l.take(sfq); pu = l.take(pu); pu0 = l.take(pu0); pxp = l.take(pxp);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_calc_work, tgrid, tblk, st, *m, tz, sfp, sfq, pu, pu0, pxp, zw, zetot,
                dt);
    return launch_status();
}

int pennant_hip_calc_work_rate(const pennant_mesh* m, const double* zvol0, const double* zvol,
                               const double* zw, const double* zp, double* zwrate,
                               const double* dt, hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_work_rate, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st,
                           m->numz, zvol0, zvol, zw, zp, zwrate, dt);
    return launch_status();
}

int pennant_hip_calc_energy(const pennant_mesh* m, const double* zetot, const double* zm,
                            double* ze, hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_energy, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st, m->numz,
                           zetot, zm, ze);
    return launch_status();
}

int pennant_hip_calc_dt_new(const pennant_mesh* m, const double* zdl, const double* zdu,
                            const double* zss, double cfl, double* dtnew, void* ws,
                            hipStream_t st) {
    if (!m || !ws || !dtnew || m->numz < 0) return PENNANT_EINVAL;
    const int nb = nred_blocks(m->numz);
    double* part = static_cast<double*>(ws);
    hipLaunchKernelGGL(k_calc_dt_new, dim3(nb), dim3(PN_BLOCK), 0, st, m->numz, zdl, zdu, zss,
                       cfl, part);
    hipLaunchKernelGGL(k_min_parts, dim3(1), dim3(PN_BLOCK), 0, st, part, nb, dtnew);
    return launch_status();
}

int pennant_hip_calc_dvol(const pennant_mesh* m, const double* zvol, const double* zvol0,
                          double* dvovmax, void* ws, hipStream_t st) {
    if (!m || !ws || !dvovmax || m->numz < 0) return PENNANT_EINVAL;
    const int nb = nred_blocks(m->numz);
    double* part = static_cast<double*>(ws);
    hipLaunchKernelGGL(k_calc_dvol, dim3(nb), dim3(PN_BLOCK), 0, st, m->numz, zvol, zvol0, part);
    hipLaunchKernelGGL(k_max_parts, dim3(1), dim3(PN_BLOCK), 0, st, part, nb, dvovmax);
    return launch_status();
}

int pennant_hip_calc_dt(const double* dtlast, const double* dtnew, const double* dvovmax,
                        double cflv, double* dtrec, hipStream_t st) {
    hipLaunchKernelGGL(k_calc_dt, dim3(1), dim3(1), 0, st, dtlast, dtnew, dvovmax, cflv, dtrec);
    return launch_status();
}

int pennant_hip_calc_global_dt(const double* dtlast, const double* dtrec, const double* time,
                               int32_t cycle, double dtmax, double dtinit, double dtfac,
                               double tstop, double* dt, hipStream_t st) {
    hipLaunchKernelGGL(k_calc_global_dt, dim3(1), dim3(1), 0, st, dtlast, dtrec, time, cycle,
                       dtmax, dtinit, dtfac, tstop, dt);
    return launch_status();
}

int pennant_hip_calc_state_half(const pennant_mesh* m, const double* zr, const double* zvolp,
                                const double* zvol0, const double* ze, const double* zwrate,
                                const double* zm, double* zp, double* zss, double gamma,
                                double ssmin, const double* dt, hipStream_t st) {
    if (!m) return PENNANT_EINVAL;
    if (m->numz > 0)
        hipLaunchKernelGGL(k_calc_state_half, dim3(nblocks(m->numz)), dim3(PN_BLOCK), 0, st,
                           m->numz, zr, zvolp, zvol0, ze, zwrate, zm, zp, zss, gamma, ssmin, dt);
    return launch_status();
}

int pennant_hip_calc_force_pgas(const pennant_mesh* m, pennant_cvec ssurf, const double* zp,
                                pennant_vec sf, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 4);
    Layout l;
    ssurf = l.take(ssurf); sf = l.take(sf);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_calc_force_pgas, tgrid, tblk, st, *m, tz, ssurf, zp, sf);
    return launch_status();
}

int pennant_hip_calc_force_tts(const pennant_mesh* m, const double* sarea, const double* smf,
                               pennant_cvec ssurf, const double* zarea, const double* zr,
                               const double* zss, pennant_vec sf, double alfa, double ssmin,
                               hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 2);
    Layout l;
    ssurf = l.take(ssurf); sf = l.take(sf);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_calc_force_tts, tgrid, tblk, st, *m, tz, sarea, smf, ssurf, zarea, zr,
                zss, sf, alfa, ssmin);
    return launch_status();
}

int pennant_hip_set_corner_div(const pennant_mesh* m, pennant_cvec ex, const double* elen,
                               pennant_cvec zx, pennant_cvec px, pennant_cvec pu,
                               pennant_vec zuc, double* carea, double* ccos, double* cdiv,
                               double* cevol, double* cdu, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 2);
    Layout l;
    ex = l.take(ex); zx = l.take(zx); px = l.take(px); pu = l.take(pu); zuc = l.take(zuc);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_set_corner_div, tgrid, tblk, st, *m, tz, ex, elen, zx, px, pu, zuc,
                carea, ccos, cdiv, cevol, cdu);
    return launch_status();
}

int pennant_hip_set_qcn_force(const pennant_mesh* m, const double* elen, const double* cdiv,
                              const double* cdu, const double* cevol, const double* zrp,
                              const double* zss, pennant_cvec pu, double* crmu,
                              pennant_vec cqe1, pennant_vec cqe2, double qgamma, double q1,
                              double q2, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 2);
    Layout l;
    pu = l.take(pu); cqe1 = l.take(cqe1); cqe2 = l.take(cqe2);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_set_qcn_force, tgrid, tblk, st, *m, tz, elen, cdiv, cdu, cevol, zrp,
                zss, pu, crmu, cqe1, cqe2, qgamma, q1, q2);
    return launch_status();
}

int pennant_hip_set_force_qcs(const pennant_mesh* m, const double* carea, pennant_cvec cqe1,
                              pennant_cvec cqe2, const double* elen, double* ccos, double* cw,
                              pennant_vec sfq, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 1);
    Layout l;
    cqe1 = l.take(cqe1); cqe2 = l.take(cqe2); sfq = l.take(sfq);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_set_force_qcs, tgrid, tblk, st, *m, tz, carea, cqe1, cqe2, elen, ccos,
                cw, sfq);
    return launch_status();
}

int pennant_hip_set_vel_diff(const pennant_mesh* m, const double* elen, const double* zss,
                             pennant_cvec px, pennant_cvec pu, double* ztmp, double* zdu,
                             double q1, double q2, hipStream_t st) {
    PN_WIDE_ENTRY(m, tz, 2);
    Layout l;
    px = l.take(px); pu = l.take(pu);
    if (l.bad) return PENNANT_EINVAL;
    PN_LAUNCH_L(l.kind(), k_set_vel_diff, tgrid, tblk, st, *m, tz, elen, zss, px, pu, ztmp, zdu,
                q1, q2);
    return launch_status();
}

int pennant_hip_apply_fixed_bc(const int32_t* mapbp, int32_t numb, double vfix_x, double vfix_y,
                               pennant_vec pf, pennant_vec pu0, hipStream_t st) {
    Layout l;
    pf = l.take(pf); pu0 = l.take(pu0);
    if (l.bad || numb < 0) return PENNANT_EINVAL;
    if (numb > 0)
        PN_LAUNCH_L(l.kind(), k_apply_fixed_bc, dim3(nblocks(numb)), dim3(PN_BLOCK), st, mapbp,
                    numb, vfix_x, vfix_y, pf, pu0);
    return launch_status();
}

}  // extern "C"
