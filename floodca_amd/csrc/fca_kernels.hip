/*
 * fca_kernels.hip -- gfx950 kernels of the WCA step (one precision per build:
 * real = float, or double with -DFCA_DOUBLE).
 *
 * FCA_HIP_FUSED: k_fused, one launch per step. Workgroup = 4 wavefronts, each
 *   walking one strip of rows over 64 columns (fca_walk.h). The workgroup
 *   first evaluates the step controller from the previous launch's record and
 *   time-step candidate, so the time step never leaves the device; workgroup 0
 *   writes the next record. Depth is double-buffered (ctrl.hbuf), hmax/tmax
 *   are updated in place by the cell's owner, the balance terms are summed per
 *   workgroup into a slot of its own (deterministic, no atomics), and the next
 *   step's candidate is one atomic min per workgroup.
 * FCA_HIP_SPLIT: the same step as controller / faces / limiter / update /
 *   candidate launches over face arrays (the phase order of the reference
 *   CUDA backend with the time step kept on the device). It shares only the
 *   scalar formulas of fca_device.h with the fused path and serves as an
 *   independent cross-check.
 */
#include <hip/hip_runtime.h>

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "fca_launch.h"
#include "fca_walk.h"

static_assert(FCA_TILE_COLS(1) == FCA_TILE_OUT(64) && FCA_TILE_COLS(2) == FCA_TILE_OUT(128),
              "host tile width differs from the walk's");

typedef real R;
typedef CandKey<R>::type Key;

#define FCA_BLOCK 256

/* ---- wavefront shifts ---------------------------------------------------- */

/* DPP wave_shl:1 / wave_shr:1: lane i reads lane i+1 / i-1; with bound_ctrl
 * lanes with no source read 0, and no prior value of the destination is
 * needed. */
__device__ __forceinline__ int lane_from_right(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ int lane_from_left(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ float lane_from_right(float x)
{
    return __int_as_float(lane_from_right(__float_as_int(x)));
}
__device__ __forceinline__ float lane_from_left(float x)
{
    return __int_as_float(lane_from_left(__float_as_int(x)));
}
__device__ __forceinline__ double lane_from_right(double x)
{
    long long b = __double_as_longlong(x);
    int lo = lane_from_right((int)(unsigned)(b & 0xffffffffll));
    int hi = lane_from_right((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double lane_from_left(double x)
{
    long long b = __double_as_longlong(x);
    int lo = lane_from_left((int)(unsigned)(b & 0xffffffffll));
    int hi = lane_from_left((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

/* Lanes (of a wavefront) whose level z + h or depth h leaves the LAND
 * form's range: |z + h| <= 2^20, |h| <= 2^14 (fca_device.h, face_land). Each
 * compare is balloted on its own, so the masks cost no vector select. */
__device__ __forceinline__ uint64_t lanes_out_of_range(float lv, float h)
{
    return __builtin_amdgcn_ballot_w64(!(__builtin_fabsf(lv) <= 0x1p20f)) |
           __builtin_amdgcn_ballot_w64(!(__builtin_fabsf(h) <= 0x1p14f));
}
__device__ __forceinline__ uint64_t lanes_out_of_range(double, double) { return 0; }

/* Lane policy of the walk on the GPU: one work-item is one lane. */
struct GpuLanes {
    static constexpr int WIDTH = 64;
    WalkLane<R> s;
    int lane;
    uint64_t outmask; /* the lanes whose column the strip writes */
    __device__ __forceinline__ void strip_begun()
    {
        outmask = __builtin_amdgcn_ballot_w64(s.outcol != 0);
    }
    /* capped faces of written columns where c holds (all lanes active): a
     * count per wavefront, kept in scalar registers, in capw */
    __device__ __forceinline__ void count_caps(WalkLane<R> &t, bool c)
    {
        t.capw += (long long)__builtin_popcountll(__builtin_amdgcn_ballot_w64(c) & outmask);
    }
    /* the work-item's share of the capped-face count */
    __device__ __forceinline__ long long cap_share() const
    {
        return s.cap + (lane == 0 ? s.capw : 0);
    }
    __device__ __forceinline__ uint64_t out_of_range_lanes()
    {
        return lanes_out_of_range(s.lv0, s.h0);
    }
    template <class F>
    __device__ __forceinline__ void each(F f)
    {
        f(s, lane);
    }
    template <class T>
    __device__ __forceinline__ void pull_right(T WalkLane<R>::*dst, T WalkLane<R>::*src)
    {
        s.*dst = lane_from_right(s.*src);
    }
    template <class T>
    __device__ __forceinline__ void pull_left(T WalkLane<R>::*dst, T WalkLane<R>::*src)
    {
        s.*dst = lane_from_left(s.*src);
    }
    __device__ __forceinline__ void count_form(int) {}
    /* Orders the wait for the loads of the row just taken (and of the masks
     * and cr/hmax fetched with them) before any later memory access: the
     * asm reads their registers, so the compiler's wait precedes it, and its
     * memory clobber keeps the following stores behind it. */
    __device__ __forceinline__ void loads_settled()
    {
        __asm__ volatile("" ::"v"(s.lv0), "v"(s.n0), "v"(s.pm1), "v"(s.pcr), "v"(s.phm)
                         : "memory");
    }
    /* uniform: the lanes whose mask field is not LAND / is LAND (a wavefront
     * is all LAND when the first is 0, has no LAND lane when the second is);
     * masks rather than bools keep the row-form bits in scalar registers */
    __device__ __forceinline__ uint64_t non_land_lanes(int WalkLane<R>::*m)
    {
        return __builtin_amdgcn_ballot_w64(s.*m != FCA_LAND);
    }
    __device__ __forceinline__ uint64_t land_lanes(int WalkLane<R>::*m)
    {
        return __builtin_amdgcn_ballot_w64(s.*m == FCA_LAND);
    }
    /* uniform: some lane's flag is set */
    __device__ __forceinline__ bool any(int WalkLane<R>::*f)
    {
        return __builtin_amdgcn_ballot_w64(s.*f != 0) != 0;
    }
    /* uniform: the largest src over the wavefront (src >= +0, not NaN) */
    __device__ __forceinline__ void wave_max_to(R WalkLane<R>::*dst, R WalkLane<R>::*src)
    {
        s.*dst = wave_max_nonneg(s.*src);
    }
    /* the work-item's share of a sum / minimum over lanes */
    template <class T>
    __device__ __forceinline__ T item_sum(T WalkLane<R>::*f)
    {
        return s.*f;
    }
    __device__ __forceinline__ R item_min(R WalkLane<R>::*f) { return s.*f; }
    __device__ __forceinline__ static float uniform(float v)
    {
        return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
    }
    __device__ __forceinline__ static double uniform(double v)
    {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffll));
        const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
        return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
    }
};

/* Lane policy with two adjacent columns per work-item: a wavefront of 64
 * work-items holds 128 lanes, lane 2i+k in s[k] of work-item i. A pull moves
 * one lane: within the work-item, or across one DPP shift between s[1] of a
 * work-item and s[0] of its neighbour. Ballots merge the two halves. The walk
 * computes the same expressions on every lane as with one column per
 * work-item; only its tiles are wider (122 written columns, against 58). */
struct GpuLanes2 {
    static constexpr int WIDTH = 128;
    WalkLane<R> s[2];
    int lane;
    uint64_t outmask[2];
    __device__ __forceinline__ void strip_begun()
    {
        outmask[0] = __builtin_amdgcn_ballot_w64(s[0].outcol != 0);
        outmask[1] = __builtin_amdgcn_ballot_w64(s[1].outcol != 0);
    }
    __device__ __forceinline__ void count_caps(WalkLane<R> &t, bool c)
    {
        const uint64_t m = (&t == &s[0]) ? outmask[0] : outmask[1];
        t.capw += (long long)__builtin_popcountll(__builtin_amdgcn_ballot_w64(c) & m);
    }
    __device__ __forceinline__ long long cap_share() const
    {
        return s[0].cap + s[1].cap + (lane == 0 ? s[0].capw + s[1].capw : 0);
    }
    __device__ __forceinline__ uint64_t out_of_range_lanes()
    {
        return lanes_out_of_range(s[0].lv0, s[0].h0) | lanes_out_of_range(s[1].lv0, s[1].h0);
    }
    template <class F>
    __device__ __forceinline__ void each(F f)
    {
        f(s[0], 2 * lane);
        f(s[1], 2 * lane + 1);
    }
    template <class T>
    __device__ __forceinline__ void pull_right(T WalkLane<R>::*dst, T WalkLane<R>::*src)
    {
        const T r = lane_from_right(s[0].*src);
        s[0].*dst = s[1].*src;
        s[1].*dst = r;
    }
    template <class T>
    __device__ __forceinline__ void pull_left(T WalkLane<R>::*dst, T WalkLane<R>::*src)
    {
        const T l = lane_from_left(s[1].*src);
        s[1].*dst = s[0].*src;
        s[0].*dst = l;
    }
    __device__ __forceinline__ void count_form(int) {}
    /* as GpuLanes::loads_settled, for both columns */
    __device__ __forceinline__ void loads_settled()
    {
        __asm__ volatile("" ::"v"(s[0].lv0), "v"(s[0].n0), "v"(s[0].pm1), "v"(s[0].pcr),
                         "v"(s[0].phm), "v"(s[1].lv0), "v"(s[1].n0), "v"(s[1].pm1),
                         "v"(s[1].pcr), "v"(s[1].phm)
                         : "memory");
    }
    /* uniform: nonzero when some lane's mask field is not LAND / is LAND (the
     * work-items whose column 0 or 1 qualifies) */
    __device__ __forceinline__ uint64_t non_land_lanes(int WalkLane<R>::*m)
    {
        return __builtin_amdgcn_ballot_w64(s[0].*m != FCA_LAND) |
               __builtin_amdgcn_ballot_w64(s[1].*m != FCA_LAND);
    }
    __device__ __forceinline__ uint64_t land_lanes(int WalkLane<R>::*m)
    {
        return __builtin_amdgcn_ballot_w64(s[0].*m == FCA_LAND) |
               __builtin_amdgcn_ballot_w64(s[1].*m == FCA_LAND);
    }
    __device__ __forceinline__ bool any(int WalkLane<R>::*f)
    {
        return (__builtin_amdgcn_ballot_w64(s[0].*f != 0) |
                __builtin_amdgcn_ballot_w64(s[1].*f != 0)) != 0;
    }
    __device__ __forceinline__ void wave_max_to(R WalkLane<R>::*dst, R WalkLane<R>::*src)
    {
        const R m = wave_max_nonneg(fmax_r(s[0].*src, s[1].*src));
        s[0].*dst = m;
        s[1].*dst = m;
    }
    template <class T>
    __device__ __forceinline__ T item_sum(T WalkLane<R>::*f)
    {
        return s[0].*f + s[1].*f;
    }
    __device__ __forceinline__ R item_min(R WalkLane<R>::*f)
    {
        return (s[1].*f < s[0].*f) ? s[1].*f : s[0].*f;
    }
};

/* The lane policy of CPL columns per work-item. */
template <int CPL>
struct LanesOf {
    typedef GpuLanes type;
};
template <>
struct LanesOf<2> {
    typedef GpuLanes2 type;
};

/* ---- reductions ----------------------------------------------------------- */

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, 64);
    }
    return v;
}
template <class T>
__device__ __forceinline__ T wave_min(T v)
{
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}

__device__ __forceinline__ void atomic_min_key(uint32_t *p, uint32_t k) { atomicMin(p, k); }
__device__ __forceinline__ void atomic_min_key(unsigned long long *p, unsigned long long k)
{
    atomicMin(p, k);
}

/* Block-wide min of a candidate, then one atomic min on the key. */
__device__ void block_min_to(R v, Key *dst)
{
    __shared__ R sh[FCA_BLOCK / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    v = wave_min(v);
    if (lane == 0) {
        sh[wave] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        R m = sh[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            m = (sh[w] < m) ? sh[w] : m;
        }
        if (m < (R)INFINITY) {
            atomic_min_key(dst, cand_key(m));
        }
    }
}

/* ---- controller record ----------------------------------------------------- */

struct CtlIO {
    const FcaDevCtrl *in;
    FcaDevCtrl *out;
    const Key *cand_in;
    Key *cand_out, *cand_reset;
    FcaHipStepRec *rec;
    int rec_cap;
};

/* Evaluates the controller; the writer (one work-item per launch) also stores
 * the next record, the step log entry and the candidate-slot bookkeeping. */
__device__ StepVals<R> ctl_step(const CtlIO &io, const FcaCtlParams &ctl, const PhysK<R> &p,
                                int writer, int pingpong, int *hbuf)
{
    FcaDevCtrl in = *io.in, nx;
    StepVals<R> sv;
    ctrl_eval<R>(in, cand_unkey(*io.cand_in), ctl, p, sv, nx);
    *hbuf = in.hbuf;
    if (pingpong && sv.active) {
        nx.hbuf = 1 - in.hbuf;
    }
    if (writer) {
        if (sv.active) {
            if (in.nrec < io.rec_cap) {
                FcaHipStepRec r;
                r.t = in.t;
                r.F = sv.F;
                r.dt_base = sv.dt_base;
                r.dt = sv.dt;
                r.step = in.step;
                r.hit_event = sv.hit;
                io.rec[in.nrec] = r;
            }
            nx.nrec = in.nrec + 1;
        } else {
            *io.cand_out = *io.cand_in; /* the state did not change */
        }
        *io.cand_reset = cand_key((R)INFINITY);
        *io.out = nx;
    }
    return sv;
}

/* ---- fused step ---------------------------------------------------------- */

struct FusedArgs {
    const R *z, *n, *cr;
    R *h0, *h1, *hmax, *tmax;
    const uint8_t *mask;
    int nr, nc, strip_rows, ngroups, ntiles, nstrips;
    PhysK<R> p;
    FcaCtlParams ctl;
    CtlIO io;
    FcaAccSlot *slots;
    /* quiet-tile skipping (fca_walk.h): flags of steps of each parity,
     * [2][ntiles], and tiles on which no cell gets rain; skip = 0 walks all */
    uint8_t *tflags;
    const uint8_t *tnorain;
    int skip;
};

/* One workgroup of the fused step: wg of nwg workgroups walking a.ntiles
 * tiles of one grid, CPL columns per work-item. */
template <int K8, int QT, int CPL>
__device__ __forceinline__ void fused_body(const FusedArgs &a, int wg, int nwg)
{
    typedef typename LanesOf<CPL>::type Lanes;
    const int outc = FCA_TILE_OUT(Lanes::WIDTH);
    __shared__ StepVals<R> sh_sv;
    __shared__ int sh_hbuf;
    __shared__ double sh_acc[FCA_BLOCK / 64][4];
    __shared__ long long sh_cap[FCA_BLOCK / 64];
    __shared__ R sh_dtc[FCA_BLOCK / 64];
    /* the wavefront index is uniform across the wavefront: readfirstlane
     * lets the walk's row loop, row addresses and bounds live in scalar
     * registers */
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int tile = wg * (FCA_BLOCK / 64) + wave;

    if (threadIdx.x == 0) {
        int hb;
        sh_sv = ctl_step(a.io, a.ctl, a.p, wg == 0, 1, &hb);
        sh_hbuf = hb;
    }
    __syncthreads();
    const StepVals<R> sv = sh_sv;
    if (!sv.active) {
        return;
    }

    Lanes lp;
    lp.lane = lane;
    walk_acc_init<R>(lp);
    {
        /* persistent wavefronts: wave w walks tiles w, w + W, w + 2W, ... in a
         * fixed order, so every workgroup's slot sums the same cells each step */
        WalkArgs<R> wa;
        const int W = nwg * (FCA_BLOCK / 64);
        wa.z = a.z, wa.n = a.n, wa.cr = a.cr, wa.mask = a.mask;
        wa.h_in = sh_hbuf ? a.h1 : a.h0;
        wa.h_out = sh_hbuf ? a.h0 : a.h1;
        wa.hmax = a.hmax, wa.tmax = a.tmax;
        wa.nr = a.nr, wa.nc = a.nc, wa.strip_rows = a.strip_rows;
        wa.p = a.p;
        wa.sv = sv;
        wa.lgg = speed_log_half_g<R>(a.p.g);
        const uint8_t *prev = a.tflags + (size_t)((sv.step + 1) & 1) * a.ntiles;
        uint8_t *cur = a.tflags + (size_t)(sv.step & 1) * a.ntiles;
        for (int t = tile; t < a.ntiles; t += W) {
            const int strip = t / a.ngroups, grp = t - strip * a.ngroups;
            const int r0 = strip * a.strip_rows;
            const int r1 = (r0 + a.strip_rows < a.nr) ? r0 + a.strip_rows : a.nr;
            if (QT && tile_skippable(prev, t, a.ngroups, a.nstrips,
                                     sv.rain_f == (R)0 || (a.tnorain[t] & 1))) {
                if (lane == 0) {
                    cur[t] = FCA_TILE_DRY | FCA_TILE_SAME;
                }
                continue;
            }
            walk_strip<R, K8, QT>(lp, wa, r0, r1, grp * outc - FCA_HALO);
            if (QT) {
                const bool wet = lp.any(&WalkLane<R>::qwet), chg = lp.any(&WalkLane<R>::qchg);
                if (lane == 0) {
                    cur[t] = (uint8_t)((wet ? 0 : FCA_TILE_DRY) | (chg ? 0 : FCA_TILE_SAME));
                }
            }
        }
    }
    walk_acc_finish<R>(lp, a.p);

    {
        typedef WalkLane<R> L;
        double vr = wave_sum(lp.item_sum(&L::v_rain)), cl = wave_sum(lp.item_sum(&L::clamp));
        double vs = wave_sum(lp.item_sum(&L::v_sea)), vb = wave_sum(lp.item_sum(&L::v_bnd));
        long long cp = wave_sum(lp.cap_share());
        R dm = wave_min(lp.item_min(&L::dtc));
        if (lane == 0) {
            sh_acc[wave][0] = vr, sh_acc[wave][1] = cl, sh_acc[wave][2] = vs;
            sh_acc[wave][3] = vb, sh_cap[wave] = cp, sh_dtc[wave] = dm;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        FcaAccSlot *sl = &a.slots[wg];
        double vr = 0, cl = 0, vs = 0, vb = 0;
        long long cp = 0;
        R dm = (R)INFINITY;
        for (int w = 0; w < FCA_BLOCK / 64; w++) {
            vr += sh_acc[w][0], cl += sh_acc[w][1], vs += sh_acc[w][2], vb += sh_acc[w][3];
            cp += sh_cap[w];
            dm = (sh_dtc[w] < dm) ? sh_dtc[w] : dm;
        }
        sl->v_rain += vr, sl->clamp += cl, sl->v_sea += vs, sl->v_bnd += vb, sl->cap += cp;
        if (sv.need_cand && dm < (R)INFINITY) {
            atomic_min_key(a.io.cand_out, cand_key(dm));
        }
    }
}

/* Wavefronts per SIMD the walk is compiled for, each the most that compiles
 * without spilling to scratch. Float, one column per work-item: the K=4 walk
 * without the quiet-tile bookkeeping fits four (at most 128 VGPRs), whose
 * fourth wavefront hides row-load latency, and so does the K=8 walk with
 * its row accesses on a scalar base (121 VGPRs; 136 with plain global
 * accesses, which ran three). The walks with the bookkeeping run three (K=4
 * 134 VGPRs, K=8 140): at four (a scalar base, or for K=4 also 8 VGPRs
 * spilled) they ran 13% (K=4) and 19% (K=8) slower on a grid 20% wet. Float, two columns per work-item: about 170 VGPRs for K=4
 * (three wavefronts, FCA_CPL2_WAVES) and more with K=8 or the bookkeeping
 * (two). Double, one column per work-item: three for K=4 (163-165 VGPRs) and
 * for K=8 without the bookkeeping, its row accesses on a scalar base (164;
 * 179 with plain global accesses, which ran two); two for K=8 with the
 * bookkeeping (184 VGPRs, plain accesses as in float). Its two-column form is
 * not built. */
#ifndef FCA_CPL2_WAVES
#define FCA_CPL2_WAVES 3
#endif
#ifdef FCA_DOUBLE
#define FCA_FUSED_WAVES(K8, QT, CPL) ((K8) && (QT) ? 2 : 3)
#else
#define FCA_FUSED_WAVES(K8, QT, CPL) \
    ((CPL) == 2 ? ((K8) || (QT) ? 2 : FCA_CPL2_WAVES) : ((QT) ? 3 : 4))
#endif
#define FCA_FUSED_OCCUPANCY(K8, QT, CPL) \
    __attribute__((amdgpu_waves_per_eu(FCA_FUSED_WAVES(K8, QT, CPL))))

template <int K8, int QT, int CPL>
__global__ void __launch_bounds__(FCA_BLOCK) FCA_FUSED_OCCUPANCY(K8, QT, CPL)
    k_fused(FusedArgs a)
{
    fused_body<K8, QT, CPL>(a, blockIdx.x, gridDim.x);
}

/* ---- batched fused step ------------------------------------------------------
 * Many grids (an ensemble of scenarios) advance one step each per launch: the
 * launch's workgroups are listed as (scenario, workgroup within it), each
 * scenario keeps its own controller, time step, candidate rotation and slots,
 * and a workgroup runs exactly the code of the single-grid launch. */

struct ScenDev {
    FusedArgs a; /* io is set per launch from ctrl/cand */
    FcaDevCtrl *ctrl;
    Key *cand;
    FcaHipStepRec *rec;
    int rec_cap, nblocks;
};

template <int K8, int CPL>
__global__ void __launch_bounds__(FCA_BLOCK) FCA_FUSED_OCCUPANCY(K8, 0, CPL)
    k_fused_batch(const ScenDev *scen, const int2 *wgs, long long k)
{
    const int2 e = wgs[blockIdx.x];
    const ScenDev *S = &scen[e.x];
    FusedArgs a = S->a;
    a.io.in = &S->ctrl[k & 1];
    a.io.out = &S->ctrl[(k + 1) & 1];
    a.io.cand_in = S->cand + (int)(k % 3);
    a.io.cand_out = S->cand + (int)((k + 1) % 3);
    a.io.cand_reset = S->cand + (int)((k + 2) % 3);
    a.io.rec = S->rec;
    a.io.rec_cap = S->rec_cap;
    fused_body<K8, 0, CPL>(a, e.y, S->nblocks);
}

/* Adds a block's balance terms to the block's own slot (a block always
 * covers the same cells, so each slot sums the same cells every step, in a
 * fixed order: the totals repeat bit for bit from run to run). */
__device__ void block_add_slot(FcaAccSlot *slot, double vr, double cl, double vs, double vb,
                               long long cap)
{
    __shared__ double sh[FCA_BLOCK / 64][4];
    __shared__ long long shc[FCA_BLOCK / 64];
    const int wave = threadIdx.x >> 6;
    vr = wave_sum(vr), cl = wave_sum(cl), vs = wave_sum(vs), vb = wave_sum(vb);
    cap = wave_sum(cap);
    if ((threadIdx.x & 63) == 0) {
        sh[wave][0] = vr, sh[wave][1] = cl, sh[wave][2] = vs, sh[wave][3] = vb;
        shc[wave] = cap;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4] = {0, 0, 0, 0};
        long long c = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            for (int j = 0; j < 4; j++) {
                t[j] += sh[w][j];
            }
            c += shc[w];
        }
        slot->v_rain += t[0], slot->clamp += t[1], slot->v_sea += t[2], slot->v_bnd += t[3];
        slot->cap += c;
    }
}

/* ---- split step ----------------------------------------------------------- */

struct SplitArgs {
    const R *z, *n, *cr;
    R *h, *hmax, *tmax, *qx, *qy, *qd1, *qd2, *qo, *s;
    const uint8_t *mask;
    int nr, nc;
    PhysK<R> p;
    FcaCtlParams ctl;
    CtlIO io;
    StepVals<R> *sv;
    FcaAccSlot *acc;
};

__global__ void k_split_ctrl(SplitArgs a)
{
    int hb;
    *a.sv = ctl_step(a.io, a.ctl, a.p, 1, 0, &hb);
}

/* Raw face fluxes (model.c:393-482): thread (r, c) owns the faces indexed
 * (r, c) in Qx, Qy, Qd1, Qd2 and the open slots of cell (r, c). */
template <int K8>
__global__ void __launch_bounds__(FCA_BLOCK) k_split_faces(SplitArgs a)
{
    const StepVals<R> sv = *a.sv;
    const int nr = a.nr, nc = a.nc;
    const size_t N = (size_t)nr * nc;
    const size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x;
    long long cap_n = 0;
    if (!sv.active) {
        return;
    }
    if (i < N) {
        const int r = (int)(i / nc), c = (int)(i - (size_t)r * nc);
        const PhysK<R> &p = a.p;
        int cap;
#define FACE(A_, B_, D_)                                                                    \
    face_flux<R>(a.z[A_], a.h[A_], a.n[A_], a.mask[A_], a.z[B_], a.h[B_], a.n[B_], a.mask[B_], \
                 (D_), (D_) == p.dx ? p.rdx : p.rddiag, (D_) == p.dx ? p.fdx : p.fddiag, sv.dt, \
                 sv.den, sv.rden, p, cap)
        if (c + 1 < nc) {
            a.qx[(size_t)r * (nc - 1) + c] = FACE(i, i + 1, p.dx);
            cap_n += cap;
        }
        if (r + 1 < nr) {
            a.qy[i] = FACE(i, i + nc, p.dx);
            cap_n += cap;
        }
        if (K8 && r + 1 < nr && c + 1 < nc) {
            a.qd1[(size_t)r * (nc - 1) + c] = FACE(i, i + nc + 1, p.ddiag);
            cap_n += cap;
            a.qd2[(size_t)r * (nc - 1) + c] = FACE(i + 1, i + nc, p.ddiag);
            cap_n += cap;
        }
#undef FACE
        if (p.open && a.mask[i] == FCA_LAND) {
            const int on[4] = {c == 0, c == nc - 1, r == 0, r == nr - 1};
            for (int k = 0; k < 4; k++) {
                if (on[k]) {
                    a.qo[4 * i + k] = open_flux<R>(a.z[i], a.h[i], a.n[i], sv.dt, sv.den, sv.rden, p, cap);
                    cap_n += cap;
                }
            }
        }
    }
    block_add_slot(a.acc + blockIdx.x, 0.0, 0.0, 0.0, 0.0, cap_n);
}

/* Outgoing volume and limiter per cell (model.c:180-238, 488-502). */
template <int K8>
__global__ void __launch_bounds__(FCA_BLOCK) k_split_limiter(SplitArgs a)
{
    const StepVals<R> sv = *a.sv;
    const int nr = a.nr, nc = a.nc;
    const size_t N = (size_t)nr * nc;
    const size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x;
    if (!sv.active || i >= N) {
        return;
    }
    const int r = (int)(i / nc), c = (int)(i - (size_t)r * nc);
    const size_t w1 = (size_t)nc - 1;
    R out = (R)0, q;
    if (c + 1 < nc && (q = a.qx[r * w1 + c]) > (R)0) {
        out += q;
    }
    if (c > 0 && (q = a.qx[r * w1 + c - 1]) < (R)0) {
        out += -q;
    }
    if (r + 1 < nr && (q = a.qy[i]) > (R)0) {
        out += q;
    }
    if (r > 0 && (q = a.qy[i - nc]) < (R)0) {
        out += -q;
    }
    if (K8) {
        if (r + 1 < nr && c + 1 < nc && (q = a.qd1[r * w1 + c]) > (R)0) {
            out += q;
        }
        if (r > 0 && c > 0 && (q = a.qd1[(r - 1) * w1 + c - 1]) < (R)0) {
            out += -q;
        }
        if (r + 1 < nr && c > 0 && (q = a.qd2[r * w1 + c - 1]) > (R)0) {
            out += q;
        }
        if (r > 0 && c + 1 < nc && (q = a.qd2[(r - 1) * w1 + c]) < (R)0) {
            out += -q;
        }
    }
    out += a.qo[4 * i] + a.qo[4 * i + 1] + a.qo[4 * i + 2] + a.qo[4 * i + 3];
    out = out * sv.dt;
    a.s[i] = (a.mask[i] == FCA_LAND) ? limiter_s<R>(out, a.h[i], a.p.A) : (R)1;
}

/* Upwind-scaled face seen from cell i; `mine` = i is the face's a side. */
__device__ __forceinline__ R scaled_face(R q, const R *s, size_t ia, size_t ib)
{
    return q * ((q >= (R)0) ? s[ia] : s[ib]);
}

/* Net inflow, rain, clamp and hmax/tmax per cell (model.c:241-319, 511-597). */
template <int K8>
__global__ void __launch_bounds__(FCA_BLOCK) k_split_update(SplitArgs a)
{
    const StepVals<R> sv = *a.sv;
    const int nr = a.nr, nc = a.nc;
    const size_t N = (size_t)nr * nc;
    const size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x;
    double vr = 0, cl = 0, vs = 0, vb = 0;
    if (!sv.active) {
        return;
    }
    if (i < N) {
        const int r = (int)(i / nc), c = (int)(i - (size_t)r * nc);
        const size_t w1 = (size_t)nc - 1;
        const R dt = sv.dt, A = a.p.A;
        const int land = a.mask[i] == FCA_LAND;
        const R *s = a.s;
        R net = (R)0, q, qos, su = s[i];
        double sea = 0.0;
#define SEA_OUT(nb_, cond_)                                              \
    if (land && a.mask[nb_] == FCA_SEA && (cond_)) {                     \
        sea += (double)dt * (double)((q > (R)0) ? q : -q);               \
    }
        if (c + 1 < nc) {
            q = scaled_face(a.qx[r * w1 + c], s, i, i + 1);
            net -= q;
            SEA_OUT(i + 1, q > (R)0)
        }
        if (c > 0) {
            q = scaled_face(a.qx[r * w1 + c - 1], s, i - 1, i);
            net += q;
            SEA_OUT(i - 1, q < (R)0)
        }
        if (r + 1 < nr) {
            q = scaled_face(a.qy[i], s, i, i + nc);
            net -= q;
            SEA_OUT(i + nc, q > (R)0)
        }
        if (r > 0) {
            q = scaled_face(a.qy[i - nc], s, i - nc, i);
            net += q;
            SEA_OUT(i - nc, q < (R)0)
        }
        if (K8) {
            if (r + 1 < nr && c + 1 < nc) {
                q = scaled_face(a.qd1[r * w1 + c], s, i, i + nc + 1);
                net -= q;
                SEA_OUT(i + nc + 1, q > (R)0)
            }
            if (r > 0 && c > 0) {
                q = scaled_face(a.qd1[(r - 1) * w1 + c - 1], s, i - nc - 1, i);
                net += q;
                SEA_OUT(i - nc - 1, q < (R)0)
            }
            if (r + 1 < nr && c > 0) {
                q = scaled_face(a.qd2[r * w1 + c - 1], s, i, i + nc - 1);
                net -= q;
                SEA_OUT(i + nc - 1, q > (R)0)
            }
            if (r > 0 && c + 1 < nc) {
                q = scaled_face(a.qd2[(r - 1) * w1 + c], s, i - nc + 1, i);
                net += q;
                SEA_OUT(i - nc + 1, q < (R)0)
            }
        }
#undef SEA_OUT
        qos = a.qo[4 * i] * su + a.qo[4 * i + 1] * su + a.qo[4 * i + 2] * su +
              a.qo[4 * i + 3] * su;
        net -= qos;
        vs = sea;
        vb = (double)dt * (double)qos;
        if (land) {
            R rain = a.cr[i] * sv.rain_f;
            R hn = a.h[i] + (dt * div_c(net, A, a.p.rA, a.p.fA) + rain);
            vr = (double)A * (double)rain;
            if (hn < (R)0) {
                cl = (double)(-hn) * (double)A;
                hn = (R)0;
            }
            a.h[i] = hn;
            if (hn > a.hmax[i]) {
                a.hmax[i] = hn;
                a.tmax[i] = sv.tmax_val;
            }
        }
    }
    block_add_slot(a.acc + blockIdx.x, vr, cl, vs, vb, 0);
}

/* Next step's candidates from the updated state (model.c:111-156). */
template <int K8>
__global__ void __launch_bounds__(FCA_BLOCK) k_split_cand(SplitArgs a)
{
    const StepVals<R> sv = *a.sv;
    const size_t N = (size_t)a.nr * a.nc;
    const size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x;
    R v = (R)INFINITY;
    if (!sv.active || !sv.need_cand) {
        return;
    }
    if (i < N) {
        const int r = (int)(i / a.nc), c = (int)(i - (size_t)r * a.nc);
        R d = cell_dt_at<R>(a.z, a.h, a.n, a.mask, r, c, a.nr, a.nc, a.p);
        if (a.mask[i] == FCA_LAND && a.h[i] >= a.p.tau_h && d == d) {
            v = d;
        }
    }
    block_min_to(v, a.io.cand_out);
}

/* ---- utility kernels ------------------------------------------------------ */

template <int K8>
__global__ void __launch_bounds__(FCA_BLOCK)
    k_cand_init(const R *z, const R *h, const R *n, const uint8_t *mask, int nr, int nc,
                PhysK<R> p, Key *dst)
{
    const size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x;
    R v = (R)INFINITY;
    if (i < (size_t)nr * nc) {
        const int r = (int)(i / nc), c = (int)(i - (size_t)r * nc);
        R d = cell_dt_at<R>(z, h, n, mask, r, c, nr, nc, p);
        if (mask[i] == FCA_LAND && h[i] >= p.tau_h && d == d) {
            v = d;
        }
    }
    block_min_to(v, dst);
}

__global__ void k_store_key(Key *dst, Key v) { *dst = v; }

/* Per tile (outc output columns): bit 0 when none of its LAND output cells
 * has cr != 0, bit 1 when it has LAND output cells. One wavefront per tile. */
__global__ void __launch_bounds__(64)
    k_tile_norain(const R *cr, const uint8_t *mask, int nr, int nc, int strip_rows, int ngroups,
                  int outc, uint8_t *out)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    const int strip = t / ngroups, grp = t - strip * ngroups;
    const int r0 = strip * strip_rows, r1 = (r0 + strip_rows < nr) ? r0 + strip_rows : nr;
    int rain = 0, land = 0;
    for (int cc = lane; cc < outc; cc += 64) {
        const int c = grp * outc + cc;
        if (c < nc) {
            for (int r = r0; r < r1; r++) {
                const size_t i = (size_t)r * nc + c;
                land |= mask[i] == FCA_LAND;
                rain |= mask[i] == FCA_LAND && !(cr[i] == (R)0);
            }
        }
    }
    rain = __builtin_amdgcn_ballot_w64(rain) != 0;
    land = __builtin_amdgcn_ballot_w64(land) != 0;
    if (lane == 0) {
        /* bit 0: no rain on the tile; bit 1: the tile holds LAND cells */
        out[t] = (uint8_t)((rain ? 0 : 1) | (land ? 2 : 0));
    }
}

__global__ void __launch_bounds__(FCA_BLOCK) k_cr(const R *C, const R *Rday, R *cr, size_t N)
{
    for (size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x; i < N;
         i += (size_t)gridDim.x * FCA_BLOCK) {
        cr[i] = C[i] * Rday[i];
    }
}

/* A crop scenario's fields from the resident base grid: z and mask copied
 * from the window at (r0, c0), n from the class of each cell, C = 1,
 * Rday = rday, and a dry start (h = hmax = tmax = 0 in both depth buffers).
 * One thread per cell of the crop, rows of ncols. */
__global__ void __launch_bounds__(FCA_BLOCK)
    k_crop(const R *bz, const uint8_t *bmask, const uint8_t *bcls, int bnc, int r0, int c0,
           int nr, int nc, FcaHipCrop crop, R *z, uint8_t *mask, R *n, R *C, R *Rday, R *h0,
           R *h1, R *hmax, R *tmax)
{
    const size_t N = (size_t)nr * nc;
    for (size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x; i < N;
         i += (size_t)gridDim.x * FCA_BLOCK) {
        const int r = (int)(i / (size_t)nc), c = (int)(i - (size_t)r * nc);
        const size_t b = (size_t)(r0 + r) * (size_t)bnc + (size_t)(c0 + c);
        z[i] = bz[b];
        mask[i] = bmask[b];
        n[i] = crop.n_class[bcls[b] & (FCA_HIP_MAX_CLASSES - 1)];
        C[i] = (R)1;
        Rday[i] = crop.rday;
        h0[i] = h1[i] = hmax[i] = tmax[i] = (R)0;
    }
}

/* Fixed-partition partial sums, so the total is the same on every run. */
__global__ void __launch_bounds__(FCA_BLOCK)
    k_volume(const R *h, const uint8_t *mask, size_t N, R A, double *red)
{
    __shared__ double sh[FCA_BLOCK / 64];
    double v = 0.0;
    for (size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x; i < N;
         i += (size_t)gridDim.x * FCA_BLOCK) {
        if (mask[i] == FCA_LAND) {
            v += (double)h[i] * (double)A;
        }
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < FCA_BLOCK / 64; w++) {
            t += sh[w];
        }
        red[blockIdx.x] = t;
    }
}

/* Completes hmax and tmax after the fused walk, which stores them only when a
 * cell stops rising: a LAND cell whose depth exceeds its stored hmax rose in
 * the last step taken, so its hmax is its depth and its tmax that step's
 * t + dt (tlast). Leaves every other cell as it is, so it may run at any
 * point between steps, any number of times. */
__global__ void __launch_bounds__(FCA_BLOCK)
    k_hmax_finalize(const R *h, R *hmax, R *tmax, const uint8_t *mask, size_t N, R tlast)
{
    for (size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x; i < N;
         i += (size_t)gridDim.x * FCA_BLOCK) {
        if (mask[i] == FCA_LAND) {
            const R v = h[i];
            if (v > hmax[i]) {
                hmax[i] = v;
                tmax[i] = tlast;
            }
        }
    }
}

/* Fills n 32-bit words with v (used where a caller might capture the work
 * into a graph, which must not hold memset nodes). */
__global__ void __launch_bounds__(FCA_BLOCK) k_fill_u32(uint32_t *p, size_t n, uint32_t v)
{
    for (size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x; i < n;
         i += (size_t)gridDim.x * FCA_BLOCK) {
        p[i] = v;
    }
}

/* Host-side checks of fca_hip_run for fields that live in device memory:
 * out[0] becomes nonzero if some LAND roughness lies outside [1e-12, 2^100]
 * (fca_land_n_ok), out[1] counts the open-boundary faces of the LAND cells
 * on the domain edge (the open part of the reference's cap_total count). */
__global__ void __launch_bounds__(FCA_BLOCK)
    k_grid_scan(const R *n, const uint8_t *mask, int nr, int nc, int open,
                unsigned long long *out)
{
    const size_t N = (size_t)nr * nc;
    unsigned long long bad = 0, faces = 0;
    for (size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x; i < N;
         i += (size_t)gridDim.x * FCA_BLOCK) {
        if (mask[i] != FCA_LAND) {
            continue;
        }
        bad |= !(n[i] >= (R)1e-12 && n[i] <= (R)0x1p100);
        if (open) {
            const int r = (int)(i / (size_t)nc), c = (int)(i - (size_t)r * nc);
            faces += (c == 0) + (c == nc - 1) + (r == 0) + (r == nr - 1);
        }
    }
    bad = wave_sum(bad);
    faces = wave_sum(faces);
    if ((threadIdx.x & 63) == 0) {
        if (bad) {
            atomicAdd(&out[0], bad);
        }
        if (faces) {
            atomicAdd(&out[1], faces);
        }
    }
}

/* ---- workspace of fca_hip_run_noalloc ---------------------------------------
 * One array per device, reserved when the library's code object is loaded and
 * held until the process ends. Host code never passes `scratch` itself to a
 * kernel (on the host the name is only a stand-in for the symbol): its device
 * address comes from hipGetSymbolAddress, once per device, with that device
 * current. Every call on a device uses the same array, so calls that overlap
 * on one device race; each call clears what it needs cleared. */
static __device__ __align__(256) unsigned char scratch[FCA_NOALLOC_SCRATCH_BYTES];

#define FCA_MAX_DEVICES 64
static void *g_scratch_ptr[FCA_MAX_DEVICES];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

extern "C" int fca_k_noalloc_scratch(void **ptr, size_t *bytes)
{
    int dev = -1, rc = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= FCA_MAX_DEVICES) {
        return -1;
    }
    pthread_mutex_lock(&g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = NULL;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            rc = -1;
        } else {
            g_scratch_ptr[dev] = p;
        }
    }
    *ptr = g_scratch_ptr[dev];
    pthread_mutex_unlock(&g_scratch_mu);
    *bytes = sizeof(scratch);
    return rc;
}

#define FCA_MAXTHR 4
struct Thr {
    double v[FCA_MAXTHR];
    int n;
};

/* Ensemble counts (ens_common.c:166-188): red holds per block
 * [flooded_0..flooded_3, max LAND hmax (at least 0)]. */
__global__ void __launch_bounds__(FCA_BLOCK)
    k_ens_count(const R *hmax, const uint8_t *mask, size_t N, Thr thr, uint32_t *counts,
                double *red)
{
    __shared__ double sh[FCA_BLOCK / 64][FCA_MAXTHR + 1];
    double f[FCA_MAXTHR] = {0, 0, 0, 0}, mx = 0.0;
    for (size_t i = (size_t)blockIdx.x * FCA_BLOCK + threadIdx.x; i < N;
         i += (size_t)gridDim.x * FCA_BLOCK) {
        if (mask[i] != FCA_LAND) {
            continue;
        }
        double hm = (double)hmax[i];
        mx = (hm > mx) ? hm : mx;
        for (int j = 0; j < thr.n; j++) {
            if (hm >= thr.v[j]) {
                if (counts) {
                    counts[(size_t)j * N + i] += 1u;
                }
                f[j] += 1.0;
            }
        }
    }
    for (int j = 0; j < FCA_MAXTHR; j++) {
        f[j] = wave_sum(f[j]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(mx, off, 64);
        mx = (o > mx) ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        for (int j = 0; j < FCA_MAXTHR; j++) {
            sh[threadIdx.x >> 6][j] = f[j];
        }
        sh[threadIdx.x >> 6][FCA_MAXTHR] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 0; j <= FCA_MAXTHR; j++) {
            double t = sh[0][j];
            for (int w = 1; w < FCA_BLOCK / 64; w++) {
                t = (j < FCA_MAXTHR) ? t + sh[w][j] : ((sh[w][j] > t) ? sh[w][j] : t);
            }
            red[(size_t)blockIdx.x * (FCA_MAXTHR + 1) + j] = t;
        }
    }
}

/* ---- launchers (C ABI) ----------------------------------------------------- */

/* PhysK of the run with the host-verified constant-divisor flags. */
static PhysK<R> phys_of(const FcaDev *d)
{
    PhysK<R> p = make_phys<R>(d->phys);
    p.fdx = d->fdiv[0], p.fddiag = d->fdiv[1], p.fA = d->fdiv[2];
    p.fast = d->land_form;
    return p;
}

static inline unsigned blocks_for(size_t n)
{
    size_t b = (n + FCA_BLOCK - 1) / FCA_BLOCK;
    return (unsigned)(b ? b : 1);
}

static inline Key *cand_slot(const FcaDev *d, long long k)
{
    return (Key *)d->cand + (int)(((k % 3) + 3) % 3);
}

static CtlIO ctl_io(const FcaDev *d, long long k)
{
    CtlIO io;
    io.in = &d->ctrl[k & 1];
    io.out = &d->ctrl[(k + 1) & 1];
    io.cand_in = cand_slot(d, k);
    io.cand_out = cand_slot(d, k + 1);
    io.cand_reset = cand_slot(d, k + 2);
    io.rec = d->rec;
    io.rec_cap = d->rec_cap;
    return io;
}

extern "C" void fca_k_geometry(FcaDev *d)
{
    if (d->strip_rows <= 0) {
        d->strip_rows = 256;
    }
    const int outc = FCA_TILE_COLS(d->cols_per_lane == 2 ? 2 : 1);
    d->ngroups = (d->nc + outc - 1) / outc;
    d->nstrips = (d->nr + d->strip_rows - 1) / d->strip_rows;
    d->ntiles = d->ngroups * d->nstrips;
    d->nblocks = (d->ntiles + FCA_BLOCK / 64 - 1) / (FCA_BLOCK / 64);
    if (d->tiles_per_wave > 1) {
        const int per_block = (FCA_BLOCK / 64) * d->tiles_per_wave;
        d->nblocks = (d->ntiles + per_block - 1) / per_block;
    }
    if (d->max_blocks > 0 && d->nblocks > d->max_blocks) {
        d->nblocks = d->max_blocks;
    }
}

extern "C" int fca_k_cand_reset(const FcaDev *d, int slot, hipStream_t st)
{
    hipLaunchKernelGGL(k_store_key, dim3(1), dim3(1), 0, st, (Key *)d->cand + slot,
                       cand_key((R)INFINITY));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fca_k_cand_init(const FcaDev *d, int hbuf, int slot, hipStream_t st)
{
    const size_t N = (size_t)d->nr * d->nc;
    const PhysK<R> p = phys_of(d);
    Key *dst = (Key *)d->cand + slot;
    if (fca_k_cand_reset(d, slot, st) != 0) {
        return -1;
    }
    if (d->K == 8) {
        hipLaunchKernelGGL(k_cand_init<1>, dim3(blocks_for(N)), dim3(FCA_BLOCK), 0, st, d->z,
                           d->h[hbuf], d->n, d->mask, d->nr, d->nc, p, dst);
    } else {
        hipLaunchKernelGGL(k_cand_init<0>, dim3(blocks_for(N)), dim3(FCA_BLOCK), 0, st, d->z,
                           d->h[hbuf], d->n, d->mask, d->nr, d->nc, p, dst);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static FusedArgs fused_args(const FcaDev *d)
{
    FusedArgs a;
    memset(&a, 0, sizeof(a));
    a.z = d->z, a.n = d->n, a.cr = d->cr, a.mask = d->mask;
    a.h0 = d->h[0], a.h1 = d->h[1], a.hmax = d->hmax, a.tmax = d->tmax;
    a.nr = d->nr, a.nc = d->nc, a.strip_rows = d->strip_rows;
    a.ngroups = d->ngroups, a.ntiles = d->ntiles, a.nstrips = d->nstrips;
    a.p = phys_of(d);
    a.ctl = d->ctl;
    a.slots = d->slots;
    a.tflags = d->tflags, a.tnorain = d->tnorain;
    a.skip = d->skip && d->tflags && d->tnorain;
    return a;
}

extern "C" int fca_k_step(const FcaDev *d, long long k, hipStream_t st)
{
    const PhysK<R> p = phys_of(d);
    if (d->kernel == FCA_HIP_FUSED) {
        FusedArgs a = fused_args(d);
        a.io = ctl_io(d, k);
        const int qt = a.skip;
        const dim3 g(d->nblocks), b(FCA_BLOCK);
#define FCA_LAUNCH_FUSED(K8_, QT_)                                                 \
    if (FCA_CPL2_BUILT && d->cols_per_lane == 2) {                                 \
        hipLaunchKernelGGL((k_fused<K8_, QT_, FCA_CPL2_BUILT ? 2 : 1>), g, b, 0, st, a); \
    } else {                                                                       \
        hipLaunchKernelGGL((k_fused<K8_, QT_, 1>), g, b, 0, st, a);                \
    }
        if (d->K == 8) {
            if (qt) {
                FCA_LAUNCH_FUSED(1, 1)
            } else {
                FCA_LAUNCH_FUSED(1, 0)
            }
        } else if (qt) {
            FCA_LAUNCH_FUSED(0, 1)
        } else {
            FCA_LAUNCH_FUSED(0, 0)
        }
#undef FCA_LAUNCH_FUSED
    } else {
        SplitArgs a;
        const size_t N = (size_t)d->nr * d->nc;
        const dim3 g(blocks_for(N)), b(FCA_BLOCK);
        a.z = d->z, a.n = d->n, a.cr = d->cr, a.mask = d->mask;
        a.h = d->h[0], a.hmax = d->hmax, a.tmax = d->tmax;
        a.qx = d->qx, a.qy = d->qy, a.qd1 = d->qd1, a.qd2 = d->qd2, a.qo = d->qo, a.s = d->s;
        a.nr = d->nr, a.nc = d->nc;
        a.p = p;
        a.ctl = d->ctl;
        a.io = ctl_io(d, k);
        a.sv = (StepVals<R> *)d->sv;
        a.acc = d->slots;
        hipLaunchKernelGGL(k_split_ctrl, dim3(1), dim3(1), 0, st, a);
        if (d->K == 8) {
            hipLaunchKernelGGL(k_split_faces<1>, g, b, 0, st, a);
            hipLaunchKernelGGL(k_split_limiter<1>, g, b, 0, st, a);
            hipLaunchKernelGGL(k_split_update<1>, g, b, 0, st, a);
            hipLaunchKernelGGL(k_split_cand<1>, g, b, 0, st, a);
        } else {
            hipLaunchKernelGGL(k_split_faces<0>, g, b, 0, st, a);
            hipLaunchKernelGGL(k_split_limiter<0>, g, b, 0, st, a);
            hipLaunchKernelGGL(k_split_update<0>, g, b, 0, st, a);
            hipLaunchKernelGGL(k_split_cand<0>, g, b, 0, st, a);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" void *fca_k_batch_table(const FcaDev *const *devs, int n)
{
    ScenDev *h = (ScenDev *)calloc((size_t)(n > 0 ? n : 1), sizeof(ScenDev));
    void *dv = NULL;
    if (!h) {
        return NULL;
    }
    for (int i = 0; i < n; i++) {
        h[i].a = fused_args(devs[i]);
        h[i].ctrl = devs[i]->ctrl;
        h[i].cand = (Key *)devs[i]->cand;
        h[i].rec = devs[i]->rec;
        h[i].rec_cap = devs[i]->rec_cap;
        h[i].nblocks = devs[i]->nblocks;
    }
    if (hipMalloc(&dv, (size_t)(n > 0 ? n : 1) * sizeof(ScenDev)) != hipSuccess ||
        hipMemcpy(dv, h, (size_t)n * sizeof(ScenDev), hipMemcpyHostToDevice) != hipSuccess) {
        if (dv) {
            (void)hipFree(dv);
        }
        dv = NULL;
    }
    free(h);
    return dv;
}

extern "C" int fca_k_step_batch(const void *tab, const void *wgs, int nwg, int K, int cpl,
                                long long k, hipStream_t st)
{
    if (nwg <= 0) {
        return 0;
    }
    const ScenDev *sc = (const ScenDev *)tab;
    const int2 *w = (const int2 *)wgs;
    const int c2 = FCA_CPL2_BUILT ? 2 : 1;
    if (K == 8) {
        if (FCA_CPL2_BUILT && cpl == 2) {
            hipLaunchKernelGGL((k_fused_batch<1, c2>), dim3(nwg), dim3(FCA_BLOCK), 0, st, sc, w, k);
        } else {
            hipLaunchKernelGGL((k_fused_batch<1, 1>), dim3(nwg), dim3(FCA_BLOCK), 0, st, sc, w, k);
        }
    } else if (FCA_CPL2_BUILT && cpl == 2) {
        hipLaunchKernelGGL((k_fused_batch<0, c2>), dim3(nwg), dim3(FCA_BLOCK), 0, st, sc, w, k);
    } else {
        hipLaunchKernelGGL((k_fused_batch<0, 1>), dim3(nwg), dim3(FCA_BLOCK), 0, st, sc, w, k);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fca_k_tile_norain(const FcaDev *d, hipStream_t st)
{
    if (!d->tnorain || d->ntiles <= 0) {
        return 0;
    }
    hipLaunchKernelGGL(k_tile_norain, dim3(d->ntiles), dim3(64), 0, st, d->cr, d->mask, d->nr,
                       d->nc, d->strip_rows, d->ngroups,
                       FCA_TILE_COLS(d->cols_per_lane == 2 ? 2 : 1), d->tnorain);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fca_k_crop(const FcaDev *d, const real *bz, const uint8_t *bmask,
                          const uint8_t *bcls, int bnc, const FcaHipCrop *crop, hipStream_t st)
{
    const size_t N = (size_t)d->nr * d->nc;
    unsigned g = blocks_for(N);
    hipLaunchKernelGGL(k_crop, dim3(g < 65536 ? g : 65536), dim3(FCA_BLOCK), 0, st, bz, bmask,
                       bcls, bnc, crop->r0, crop->c0, d->nr, d->nc, *crop, d->z, d->mask, d->n,
                       d->C, d->Rday, d->h[0], d->h[1], d->hmax, d->tmax);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fca_k_cr(const FcaDev *d, hipStream_t st)
{
    const size_t N = (size_t)d->nr * d->nc;
    unsigned g = blocks_for(N);
    hipLaunchKernelGGL(k_cr, dim3(g < 65536 ? g : 65536), dim3(FCA_BLOCK), 0, st, d->C, d->Rday,
                       d->cr, N);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fca_k_volume(const FcaDev *d, int hbuf, hipStream_t st)
{
    const size_t N = (size_t)d->nr * d->nc;
    unsigned g = blocks_for(N);
    if (g > (unsigned)d->nred) {
        g = (unsigned)d->nred;
    }
    hipLaunchKernelGGL(k_volume, dim3(g), dim3(FCA_BLOCK), 0, st, d->h[hbuf], d->mask, N,
                       (R)d->phys.A, d->red);
    return hipGetLastError() == hipSuccess ? (int)g : -1;
}

extern "C" int fca_k_ens_count(const FcaDev *d, uint32_t *counts, int nthr,
                               const double *thr_host, hipStream_t st)
{
    const size_t N = (size_t)d->nr * d->nc;
    unsigned g = blocks_for(N);
    Thr t;
    if (nthr < 1 || nthr > FCA_MAXTHR) {
        return -1;
    }
    for (int j = 0; j < FCA_MAXTHR; j++) {
        t.v[j] = j < nthr ? thr_host[j] : 0.0;
    }
    t.n = nthr;
    if (g > (unsigned)(d->nred / (FCA_MAXTHR + 1))) {
        g = (unsigned)(d->nred / (FCA_MAXTHR + 1));
    }
    hipLaunchKernelGGL(k_ens_count, dim3(g), dim3(FCA_BLOCK), 0, st, d->hmax, d->mask, N, t,
                       counts, d->red);
    return hipGetLastError() == hipSuccess ? (int)g : -1;
}

extern "C" int fca_k_zero(void *p, size_t bytes, hipStream_t st)
{
    const size_t n = (bytes + 3) / 4;
    unsigned g = blocks_for(n);
    if (n == 0) {
        return 0;
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(g < 4096 ? g : 4096), dim3(FCA_BLOCK), 0, st,
                       (uint32_t *)p, n, 0u);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fca_k_grid_scan(const FcaDev *d, unsigned long long *out, hipStream_t st)
{
    const size_t N = (size_t)d->nr * d->nc;
    unsigned g = blocks_for(N);
    if (fca_k_zero(out, 2 * sizeof(unsigned long long), st) != 0) {
        return -1;
    }
    hipLaunchKernelGGL(k_grid_scan, dim3(g < 4096 ? g : 4096), dim3(FCA_BLOCK), 0, st, d->n,
                       d->mask, d->nr, d->nc, d->phys.boundary == FCA_BND_OPEN, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fca_k_hmax_finalize(const FcaDev *d, int hbuf, double tlast, hipStream_t st)
{
    const size_t N = (size_t)d->nr * d->nc;
    unsigned g = blocks_for(N);
    hipLaunchKernelGGL(k_hmax_finalize, dim3(g < 65536 ? g : 65536), dim3(FCA_BLOCK), 0, st,
                       d->h[hbuf], d->hmax, d->tmax, d->mask, N, (R)tlast);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
