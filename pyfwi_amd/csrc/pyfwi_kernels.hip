// Stencil, injection and correlation kernels of the elastic propagator for
// gfx950 (MI355X).  FP32 throughout, no MFMA: the path is HBM-bandwidth bound.
//
// Device layout (see pyfwi_plan.hip): every field and model array is a padded
// row-major grid with row pitch P (a multiple of 64 floats, P >= tnx + 8) and
// HALO zero rows above and below; cell (i, j) of the reference's tnz x tnx
// buffer lives at origin + i*P + j, with origin = HALO*P + 64 - sdo so that
// column sdo starts a 256-byte line.  The P - tnx zero columns after a row
// double as the left halo of the next row, so every stencil tap of an active cell reads
// either a real cell or a zero, which is what the reference's flat-index taps
// see inside its never-updated outer ring (src/PyFWI/elastic.cl:163-166).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/pyfwi_hip.h"
#include "pyfwi_kernels.h"

namespace pyfwi {

// ------------------------------------------------------------------ helpers

// Staggered derivatives at p along stride s with pre-divided coefficients
// c[k] = c_k / h.  Minus: pairs (0,-1) (1,-2) (2,-3) (3,-4); plus: pairs
// (1,0) (2,-1) (3,-2) (4,-3) -- the reference's Dx-/Dx+/Dz-/Dz+
// (src/PyFWI/elastic.cl:168-172, 208-213).
template <int SDO>
__device__ __forceinline__ float dminus(const float *__restrict__ p, int s,
                                        const float c[4])
{
    float r = c[0] * (p[0] - p[-s]);
    r = fmaf(c[1], p[s] - p[-2 * s], r);
    if constexpr (SDO == 4) {
        r = fmaf(c[2], p[2 * s] - p[-3 * s], r);
        r = fmaf(c[3], p[3 * s] - p[-4 * s], r);
    }
    return r;
}

template <int SDO>
__device__ __forceinline__ float dplus(const float *__restrict__ p, int s,
                                       const float c[4])
{
    float r = c[0] * (p[s] - p[0]);
    r = fmaf(c[1], p[2 * s] - p[-s], r);
    if constexpr (SDO == 4) {
        r = fmaf(c[2], p[3 * s] - p[-2 * s], r);
        r = fmaf(c[3], p[4 * s] - p[-3 * s], r);
    }
    return r;
}

// Active shot tests shared by the stencil kernels.  Forward steps run for
// t < nt; backward steps (reconstruction and adjoint) for 1 <= t < nt, the
// reference's loop `for t in range(nt-1, 0, -1)`.

__device__ __forceinline__ bool fwd_active(const pyfwi_shot *shots, int s, int t)
{
    return t < shots[s].nt;
}

__device__ __forceinline__ bool bwd_active(const pyfwi_shot *shots, int s, int t)
{
    return t >= 1 && t < shots[s].nt;
}

// ------------------------------------------------------ column marching
//
// The stencil kernels below give each thread one column j and a strip of
// rows (height chosen per launch, see strip_rows).  The thread walks down the
// strip keeping, for every field differentiated along z, a register window of
// the 2*SDO+1 rows around the current row, so each row of such a field is
// fetched from memory once.
// Derivatives along x read the neighbouring columns of the same row; those
// loads are coalesced across the wavefront and served by L1 / L2.

template <int SDO>
struct ZWin {
    float v[2 * SDO + 1];  // v[SDO + k] = row r + k
    __device__ __forceinline__ void push(float x)
    {
#pragma unroll
        for (int k = 0; k < 2 * SDO; ++k) v[k] = v[k + 1];
        v[2 * SDO] = x;
    }
    __device__ __forceinline__ float centre() const { return v[SDO]; }
    // backward-staggered: pairs (0,-1) (1,-2) (2,-3) (3,-4)
    __device__ __forceinline__ float dminus(const float c[4]) const
    {
        float r = c[0] * (v[SDO] - v[SDO - 1]);
        r = fmaf(c[1], v[SDO + 1] - v[SDO - 2], r);
        if constexpr (SDO == 4) {
            r = fmaf(c[2], v[SDO + 2] - v[SDO - 3], r);
            r = fmaf(c[3], v[SDO + 3] - v[SDO - 4], r);
        }
        return r;
    }
    // forward-staggered: pairs (1,0) (2,-1) (3,-2) (4,-3)
    __device__ __forceinline__ float dplus(const float c[4]) const
    {
        float r = c[0] * (v[SDO + 1] - v[SDO]);
        r = fmaf(c[1], v[SDO + 2] - v[SDO - 1], r);
        if constexpr (SDO == 4) {
            r = fmaf(c[2], v[SDO + 3] - v[SDO - 2], r);
            r = fmaf(c[3], v[SDO + 4] - v[SDO - 3], r);
        }
        return r;
    }
};

// A field seen through a byte offset: `base` is wave-uniform (scalar
// registers) and the per-lane part is one 32-bit byte offset shared by all
// fields with the same layout, so each extra field costs no vector registers
// for its address.  at(o, k) reads k columns to the right of offset o.
struct CellRef {
    const float *base;
    __device__ __forceinline__ float operator()(uint32_t o) const
    {
        return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + o);
    }
    __device__ __forceinline__ float at(uint32_t o, int k) const
    {
        return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base + k) + o);
    }
};

__device__ __forceinline__ float &cell_ref(float *base, uint32_t o)
{
    return *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + o);
}

// Staggered x-derivatives of a quantity given per column offset k.
template <int SDO, class V>
__device__ __forceinline__ float xminus(V v, const float c[4])
{
    float r = c[0] * (v(0) - v(-1));
    r = fmaf(c[1], v(1) - v(-2), r);
    if constexpr (SDO == 4) {
        r = fmaf(c[2], v(2) - v(-3), r);
        r = fmaf(c[3], v(3) - v(-4), r);
    }
    return r;
}

template <int SDO, class V>
__device__ __forceinline__ float xplus(V v, const float c[4])
{
    float r = c[0] * (v(1) - v(0));
    r = fmaf(c[1], v(2) - v(-1), r);
    if constexpr (SDO == 4) {
        r = fmaf(c[2], v(3) - v(-2), r);
        r = fmaf(c[3], v(4) - v(-3), r);
    }
    return r;
}

// The 2*SDO values a staggered x-derivative reads, loaded into registers:
// t[k] = field at column offset k - SDO + 1 - M, with M = 1 for the
// backward-staggered (xminus, offsets -SDO..SDO-1) and M = 0 for the
// forward-staggered form (xplus, offsets -SDO+1..SDO).  Loading the taps
// apart from the arithmetic lets a conditional load be issued with the
// other loads of a row instead of waiting inside its branch.
template <int SDO, int M>
struct Taps {
    float t[2 * SDO] = {};
    __device__ __forceinline__ void load(const CellRef &f, uint32_t o)
    {
#pragma unroll
        for (int k = 0; k < 2 * SDO; ++k) t[k] = f.at(o, k - SDO + 1 - M);
    }
    __device__ __forceinline__ float operator()(int k) const { return t[k + SDO - 1 + M]; }
};

// Delay line of N values: pop() returns the oldest value and appends the
// newest, so a value comes back N calls after it went in.  The marching
// kernels take the x-derivatives of a windowed field on the row entering its
// z-window, while that row's cache lines are fresh, and delay them SDO rows
// to the centre row where they are used; taking them at the centre row
// instead re-fetches the row from memory.
template <int N>
struct Fifo {
    float v[N];
    __device__ __forceinline__ float pop(float x)
    {
        const float out = v[0];
#pragma unroll
        for (int k = 0; k < N - 1; ++k) v[k] = v[k + 1];
        v[N - 1] = x;
        return out;
    }
};

// ------------------------------------------------ x-taps from lane shifts
//
// At order 8 a staggered x-derivative reads 8 columns of a row; loading each
// tap costs the adjoint kernels 8 loads and registers per array.  Instead a
// lane loads its own column and one halo value, and the taps come from the
// neighbouring lanes through DPP moves (GFX9 encodings: a wave shift by one
// lane, a row shift by n lanes within the 16-lane rows; a lane whose source
// is out of range keeps the `old` operand).  Every lane of the wave must be
// active at these moves.
constexpr int DPP_ROW_SHL = 0x100, DPP_ROW_SHR = 0x110;      // + n, 1 <= n <= 15
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;  // lane i <- i+1, i <- i-1

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float old, float v)
{
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Byte offset of a lane's halo column: lanes 0..SDO-1 load columns j0-SDO ..
// j0-1 and lanes 64-SDO..63 columns j0+64 .. j0+63+SDO (j0: lane 0's column);
// the others load their own column again.
template <int SDO>
__device__ __forceinline__ uint32_t halo_offset(int lane)
{
    const int k = lane < SDO ? -SDO : (lane >= 64 - SDO ? SDO : 0);
    return (uint32_t)(k * 4);
}

// The halo value that fills the wave's edge lanes at distance k: lane 0 gets
// column j0-k (halo lane SDO-k), lane 63 gets column j0+63+k (halo lane
// 63-SDO+k).
template <int SDO, int K>
__device__ __forceinline__ float halo_fill(float h)
{
    constexpr int N = SDO - K;
    if constexpr (N == 0) {
        return h;
    } else {
        const float right = dpp_mov<DPP_ROW_SHR + N>(h, h);  // lane 63 <- lane 63-N
        return dpp_mov<DPP_ROW_SHL + N>(right, h);           // lane 0 <- lane N
    }
}

// A quantity at column offsets -SDO..SDO of the lane, from its own value `c`
// and its halo value `h` (see halo_offset); usable with xminus / xplus.
template <int SDO>
struct LaneRow {
    float s[2 * SDO + 1];
    __device__ __forceinline__ LaneRow(float c, float h)
    {
        s[SDO] = c;
        fill<1>(c, c, h);
    }
    template <int K>
    __device__ __forceinline__ void fill(float l, float r, float h)
    {
        const float e = halo_fill<SDO, K>(h);
        l = dpp_mov<DPP_WAVE_SHR1>(e, l);
        r = dpp_mov<DPP_WAVE_SHL1>(e, r);
        s[SDO - K] = l;
        s[SDO + K] = r;
        if constexpr (K < SDO) fill<K + 1>(l, r, h);
    }
    __device__ __forceinline__ float operator()(int k) const { return s[SDO + k]; }
};

// Index of the calling wave in its block, in a scalar register: the compiler
// cannot tell on its own that threadIdx.x / 64 is the same for a whole wave,
// and would otherwise keep the shot index and every pointer derived from it
// in vector registers.
__device__ __forceinline__ int wave_id()
{
    return __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64);
}

// Tile of this block and shot / column / row strip of this thread.
//
// A block is 64*SPB threads: SPB shots side by side, one wave each, over the
// same 64 columns of one row strip (a narrow grid leaves few lanes idle in
// its last column block).  Launches are 1-D.  The tiles (shot group, column
// block, row strip) are numbered shot group fastest; workgroup w runs tile
// (w % 8) * ceil(T/8) + w / 8.  Workgroups are handed to the 8 XCDs in turn,
// so each XCD receives one contiguous run of tiles: the same cells of every
// shot of the launch meet in one L2, and neighbouring column blocks, which
// share halo columns, run side by side.  A launch of one shot (which shares
// no accumulators) gives the block's SPB waves SPB adjacent column blocks of
// one strip instead, so no wave idles.  lane_ok: the lane's column is on the
// grid; the result: the wave has a tile.
template <int SDO, int SPB>
__device__ __forceinline__ bool tile_lane(const DevGrid &g, int &s, int &j, int &i0, int &i1,
                                          bool &lane_ok)
{
    if (SPB > 1 && g.nshot == 1) {
        // one-shot launch: tiles of SPB column blocks, one per wave
        const int ncq = (g.ncb + SPB - 1) / SPB;
        const int T = ncq * g.nstrips;
        const int per = (T + NUM_XCD - 1) / NUM_XCD;
        const int w = (int)blockIdx.x;
        const int logical = (w % NUM_XCD) * per + w / NUM_XCD;
        if (logical >= T) return false;
        const int cb = (logical % ncq) * SPB + wave_id();
        const int st = logical / ncq;
        s = 0;
        j = SDO + cb * 64 + (int)threadIdx.x % 64;
        i0 = SDO + st * g.mrows;
        i1 = min(i0 + g.mrows - 1, g.nz - SDO);
        lane_ok = j <= g.nx - SDO;
        return cb < g.ncb;
    }
    const int T = g.ncb * g.nstrips * g.nbatch;
    const int per = (T + NUM_XCD - 1) / NUM_XCD;
    const int w = (int)blockIdx.x;
    const int logical = (w % NUM_XCD) * per + w / NUM_XCD;
    if (logical >= T) return false;
    const int grp = logical % g.nbatch;
    const int rest = logical / g.nbatch;
    const int cb = rest % g.ncb;
    const int st = rest / g.ncb;
    const int wave = wave_id(), lane = (int)threadIdx.x % 64;
    s = grp * SPB + wave;
    j = SDO + cb * 64 + lane;
    i0 = SDO + st * g.mrows;
    i1 = min(i0 + g.mrows - 1, g.nz - SDO);
    lane_ok = j <= g.nx - SDO;
    return s < g.nshot;
}

// As tile_lane, false also for a lane whose column is off the grid.
template <int SDO, int SPB>
__device__ __forceinline__ bool tile(const DevGrid &g, int &s, int &j, int &i0, int &i1)
{
    bool lane_ok = false;
    return tile_lane<SDO, SPB>(g, s, j, i0, i1, lane_ok) && lane_ok;
}

// Prime a window for the first row i0 of a strip: rows i0-SDO .. i0+SDO-1.
// `val(k)` returns the windowed quantity at padded offset k.
template <int SDO, class F>
__device__ __forceinline__ void prime(ZWin<SDO> &w, int64_t c0, int pitch, F val)
{
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) w.v[k + 1] = val(c0 + (int64_t)(k - SDO) * pitch);
}

// ------------------------------------------- sources, records, residuals
//
// The forward stencil kernels apply the sparse terms themselves instead of a
// separate launch: the source of step t is added to the values they load at
// the source cell (vx, vz, txx, tzz), and the record sample of step t+1 is
// written by the stress update of step t at the receiver cells.  The sums are
// the float additions the reference makes in memory (elastic_surface.cl:
// 7-41), so the results are the same; the state in memory itself stays free
// of source terms.  The adjoint's residual enters memory as in the reference,
// but written by the adjoint kernels themselves: the adjoint stresses by the
// stress update of the step, the adjoint velocities by the velocity update of
// the step before (in time reversal), after it used them for its own imaging
// term; only a shot's first adjoint step needs a launch of its own.

// Source of step t for one shot (injSrc, elastic_surface.cl:34-40); z = -1
// when the step injects nothing.
struct SrcTerm {
    int z = -1, x = -1;
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
};

__device__ __forceinline__ SrcTerm src_term(const pyfwi_shot &sh, int t, const float *src)
{
    SrcTerm r;
    if (src && t < sh.nt && (t % sh.dt_scale) == 0) {
        const float *sv = src + sh.src_offset + (int64_t)(t / sh.dt_scale) * 5;
        r.z = sh.src_z;
        r.x = sh.src_x;
#pragma unroll
        for (int f = 0; f < 4; ++f) r.v[f] = sv[f];
    }
    return r;
}

// A receiver line of one shot at one time sample: `base` points at component
// 0, receiver 0 of the sample's row; component f is `fs` floats further.  z
// = -1 when nothing is recorded or injected at this step.
struct RecLine {
    int z = -1, x0 = 0, dx = 1, n = 0;
    int64_t fs = 0;
    const float *base = nullptr;
    // receiver index of column `col` on the line, or -1
    __device__ __forceinline__ int index(int col) const
    {
        const int d = col - x0;
        if (d < 0 || d % dx != 0) return -1;
        const int r = d / dx;
        return r < n ? r : -1;
    }
    __device__ __forceinline__ float value(int col, int f) const
    {
        const int r = index(col);
        return r >= 0 ? base[f * fs + r] : 0.0f;
    }
};

// Sample of step t1 = t + 1 written by the forward step t (the reference
// records before step t1's injection, which is the state after step t).
__device__ __forceinline__ RecLine record_line(const pyfwi_shot &sh, int t1, float *seis)
{
    RecLine q;
    if (!seis || t1 >= sh.nt || (t1 % sh.dt_scale) != 0) return q;
    const int k = t1 / sh.dt_scale;
    if (k >= sh.n_samples) return q;
    q.z = sh.rec_z;
    q.x0 = sh.rec_x0;
    q.dx = sh.rec_dx;
    q.n = sh.n_rec;
    q.fs = (int64_t)sh.n_samples * sh.rec_stride;
    q.base = seis + sh.rec_offset + (int64_t)k * sh.rec_stride;
    return q;
}

// Residual row injected at backward step t (Adj_injSrc, elastic_surface.cl:
// 48-72; wave_propagation.py:815-824): the reference counts injections down
// from nt_ext, so step t uses row nt_ext-1-(floor((nt-1)/dt_scale) -
// t/dt_scale), a negative row wrapping like a Python index.  z = -1 when the
// step injects nothing.
__device__ __forceinline__ RecLine residual_line(const pyfwi_shot &sh, int t, const float *res)
{
    RecLine q;
    if (!res || t < 1 || t >= sh.nt || (t % sh.dt_scale) != 0) return q;
    int row = sh.n_samples - 1 - ((sh.nt - 1) / sh.dt_scale - t / sh.dt_scale);
    if (row < 0) row += sh.n_samples;
    if (row < 0) return q;
    q.z = sh.rec_z;
    q.x0 = sh.rec_x0;
    q.dx = sh.rec_dx;
    q.n = sh.n_rec;
    q.fs = (int64_t)sh.n_samples * sh.rec_stride;
    q.base = res + sh.rec_offset + (int64_t)row * sh.rec_stride;
    return q;
}

// ---------------------------------------------------------- forward sweeps

// Velocity half step of every active shot (src/PyFWI/elastic.cl:155-226).
//   DIR = +1: v += dt*b*D(tau); v -= dt*(dx+dz)*v   (damping after)
//   DIR = -1: v -= dt*(dx+dz)*v; v -= dt*b*D(tau)   (time reversed)
//
// The state is read from `in` and the updated velocities written to `out`
// (per-shot strides in_stride / out_stride, in floats).  in == out updates in
// place; the forward of a kept-state run reads the state of step t-1 from its
// slot and writes step t into the next slot.
template <int SDO, int DIR, int SPB>
__global__ __launch_bounds__(64 * SPB) void k_vel(DevGrid g, const pyfwi_shot *shots, int t,
                                                 const float *in, int64_t in_stride, float *out,
                                                 int64_t out_stride, const float *__restrict__ bu,
                                                 const float *__restrict__ dxp,
                                                 const float *__restrict__ dzp,
                                                 const float *__restrict__ src)
{
    int s, j, i0, i1;
    if (!tile<SDO, SPB>(g, s, j, i0, i1)) return;
    if (DIR > 0 ? !fwd_active(shots, s, t) : !bwd_active(shots, s, t)) return;
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    const float *F = in + (int64_t)s * in_stride + g.origin;
    float *O = out + (int64_t)s * out_stride + g.origin;
    const CellRef vx{F}, vz{F + fs}, txx{F + 2 * fs}, tzz{F + 3 * fs}, txz{F + 4 * fs};
    const CellRef B{bu + g.origin};
    // the step's source term (forward only): vx, vz at the centre, txx in the
    // x-taps and tzz in the z-window at the source cell.  The terms this
    // thread adds are fixed here; the loop only adds registers.
    SrcTerm st = DIR > 0 ? src_term(shots[s], t, src) : SrcTerm{};
    if (st.z < i0 - SDO || st.z > i1 + SDO || st.x < j - SDO || st.x > j + SDO) st.z = -1;
    float s_txx[2 * SDO + 1];  // txx term at tap offset k (index k + SDO)
#pragma unroll
    for (int k = -SDO; k <= SDO; ++k) s_txx[k + SDO] = j + k == st.x ? st.v[2] : 0.0f;
    const bool src_col = j == st.x;
    const float s_vx = src_col ? st.v[0] : 0.0f, s_vz = src_col ? st.v[1] : 0.0f;
    const float s_tzz = src_col ? st.v[3] : 0.0f;
    auto tzz_row = [&](uint32_t o, int row) {
        float v = tzz(o);
        if (row == st.z) v += s_tzz;
        return v;
    };
    ZWin<SDO> wxz, wzz;  // txz for d/dz-, tzz for d/dz+
    Fifo<SDO> fxz;       // d/dx+ txz
    auto dx_txz = [&](uint32_t o) { return xplus<SDO>([&](int k) { return txz.at(o, k); }, g.cx); };
    uint32_t cb = ((uint32_t)i0 * (uint32_t)g.pitch + (uint32_t)j) * 4u;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        wxz.v[k + 1] = txz(o);
        wzz.v[k + 1] = tzz_row(o, i0 + k - SDO);
        if (k >= SDO) fxz.v[k - SDO] = dx_txz(o);
    }
    const float ddx = dxp[j];
#pragma unroll MARCH_UNROLL
    for (int i = i0; i <= i1; ++i, cb += RB) {
        const uint32_t n = cb + (uint32_t)SDO * RB;
        wxz.push(txz(n));
        wzz.push(tzz_row(n, i + SDO));
        const float dxp_txz = fxz.pop(dx_txz(n));
        const float bc = B(cb);
        Taps<SDO, 1> ttxx;  // loaded outside the branch below, which then holds no load
        ttxx.load(txx, cb);
        float ux = vx(cb), uz = vz(cb);
        if (i == st.z) {  // the source row
#pragma unroll
            for (int k = 0; k < 2 * SDO; ++k) ttxx.t[k] += s_txx[k];
            ux += s_vx;
            uz += s_vz;
        }
        const float dxm_txx = xminus<SDO>(ttxx, g.cx);
        const float ax = bc * (dxm_txx + wxz.dminus(g.cz));
        const float az = bc * (wzz.dplus(g.cz) + dxp_txz);
        const float damp = ddx + dzp[i];
        if (DIR > 0) {
            ux = fmaf(g.dt, ax, ux);
            ux = fmaf(g.dt, -damp * ux, ux);
            uz = fmaf(g.dt, az, uz);
            uz = fmaf(g.dt, -damp * uz, uz);
        } else {
            ux = fmaf(-g.dt, damp * ux, ux);
            ux = fmaf(-g.dt, ax, ux);
            uz = fmaf(-g.dt, damp * uz, uz);
            uz = fmaf(-g.dt, az, uz);
        }
        cell_ref(O, cb) = ux;
        cell_ref(O + fs, cb) = uz;
    }
}

// Stress half step (src/PyFWI/elastic.cl:235-308).  Forward: the velocities
// of this step are read from `out` (written there by k_vel), the stresses of
// the previous step from `in`, and the new stresses written to `out`.  Time
// reversed: in == out.
template <int SDO, int DIR, int SPB>
__global__ __launch_bounds__(64 * SPB) void k_str(DevGrid g, const pyfwi_shot *shots, int t,
                                                 const float *in, int64_t in_stride, float *out,
                                                 int64_t out_stride, const float *__restrict__ lam,
                                                 const float *__restrict__ mu,
                                                 const float *__restrict__ dxp,
                                                 const float *__restrict__ dzp,
                                                 const float *__restrict__ src, float *seis)
{
    int s, j, i0, i1;
    if (!tile<SDO, SPB>(g, s, j, i0, i1)) return;
    if (DIR > 0 ? !fwd_active(shots, s, t) : !bwd_active(shots, s, t)) return;
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    const float *F = in + (int64_t)s * in_stride + g.origin;
    float *O = out + (int64_t)s * out_stride + g.origin;
    const CellRef vx{O}, vz{O + fs}, txx{F + 2 * fs}, tzz{F + 3 * fs}, txz{F + 4 * fs};
    const CellRef L{lam + g.origin}, M{mu + g.origin};
    // forward only: the source term on txx, tzz at the source cell, and the
    // record sample of step t+1 (this step's result) at the receivers
    SrcTerm st = DIR > 0 ? src_term(shots[s], t, src) : SrcTerm{};
    if (st.x != j || st.z < i0 || st.z > i1) st.z = -1;
    RecLine rl = DIR > 0 ? record_line(shots[s], t + 1, seis) : RecLine{};
    const int rec_r = rl.z >= i0 && rl.z <= i1 ? rl.index(j) : -1;
    if (rec_r < 0) rl.z = -1;
    ZWin<SDO> wvz, wvx;  // vz for d/dz-, vx for d/dz+
    Fifo<SDO> fxx, fxz;  // d/dx+ vx, d/dx- vz
    auto dx_vx = [&](uint32_t o) { return xplus<SDO>([&](int k) { return vx.at(o, k); }, g.cx); };
    auto dx_vz = [&](uint32_t o) { return xminus<SDO>([&](int k) { return vz.at(o, k); }, g.cx); };
    uint32_t cb = ((uint32_t)i0 * (uint32_t)g.pitch + (uint32_t)j) * 4u;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        wvz.v[k + 1] = vz(o);
        wvx.v[k + 1] = vx(o);
        if (k >= SDO) {
            fxx.v[k - SDO] = dx_vx(o);
            fxz.v[k - SDO] = dx_vz(o);
        }
    }
    const float ddx = dxp[j];
#pragma unroll MARCH_UNROLL
    for (int i = i0; i <= i1; ++i, cb += RB) {
        const uint32_t n = cb + (uint32_t)SDO * RB;
        wvz.push(vz(n));
        wvx.push(vx(n));
        const float exx = fxx.pop(dx_vx(n));
        const float dxm_vz = fxz.pop(dx_vz(n));
        const float ezz = wvz.dminus(g.cz);
        const float gxz = dxm_vz + wvx.dplus(g.cz);
        const float l = L(cb), m = M(cb);
        const float lm = l + 2.0f * m;
        // the fused multiply-adds written out, so that the one-launch step
        // (k_fwd_step) rounds as this kernel does
        const float sxx = fmaf(l, ezz, lm * exx);
        const float szz = fmaf(l, exx, lm * ezz);
        const float sxz = m * gxz;
        const float damp = ddx + dzp[i];
        float a = txx(cb), bb = tzz(cb), d = txz(cb);
        if (i == st.z) {
            a += st.v[2];
            bb += st.v[3];
        }
        if (DIR > 0) {
            a = fmaf(g.dt, sxx, a);
            bb = fmaf(g.dt, szz, bb);
            d = fmaf(g.dt, sxz, d);
            a = fmaf(g.dt, -damp * a, a);
            bb = fmaf(g.dt, -damp * bb, bb);
            d = fmaf(g.dt, -damp * d, d);
        } else {
            a = fmaf(-g.dt, damp * a, a);
            bb = fmaf(-g.dt, damp * bb, bb);
            d = fmaf(-g.dt, damp * d, d);
            a = fmaf(-g.dt, sxx, a);
            bb = fmaf(-g.dt, szz, bb);
            d = fmaf(-g.dt, sxz, d);
        }
        cell_ref(O + 2 * fs, cb) = a;
        cell_ref(O + 3 * fs, cb) = bb;
        cell_ref(O + 4 * fs, cb) = d;
        if (i == rl.z) {  // this thread's receiver
            float *w = const_cast<float *>(rl.base) + rec_r;
            w[0] = wvx.centre();
            w[rl.fs] = wvz.centre();
            w[2 * rl.fs] = a;
            w[3 * rl.fs] = bb;
            w[4 * rl.fs] = d;
        }
    }
}

// ------------------------------------------------ fused two-stage steps
//
// One launch performs both half steps of a time step, reading the state from
// `in` and writing the new state to `out`, so every field is read once and
// written once (52 B per cell instead of 72 B for two launches).
//
// Stage 1 updates field group X at row r; stage 2 updates group Y at row
// q = r - SDO from the NEW X values, whose z-neighbours sit in a register
// window and whose x-neighbours come from the adjacent lanes by shuffle.
// A wavefront covers 64 columns; its SDO lanes on each side only recompute
// stage-1 values for their neighbours, the 64 - 2*SDO lanes in between own
// their column and write.  Each strip also recomputes stage 1 on the SDO rows
// above and below it.  Cells outside the active range stay zero.

template <int SDO>
struct FusedTile {
    int s, col, i0, i1;
    bool own, colact;
};

template <int SDO>
__device__ __forceinline__ bool fused_tile(const DevGrid &g, FusedTile<SDO> &ft)
{
    const int T = g.ncb * g.nstrips * g.nbatch;
    const int per = (T + NUM_XCD - 1) / NUM_XCD;
    const int w = (int)blockIdx.x;
    const int logical = (w % NUM_XCD) * per + w / NUM_XCD;
    if (logical >= T) return false;
    const int cb = logical % g.ncb;
    const int rest = logical / g.ncb;
    const int st = rest % g.nstrips;
    ft.s = rest / g.nstrips;
    const int wave = wave_id(), lane = (int)threadIdx.x % 64;
    const int span = 64 - 2 * SDO;
    const int x0 = SDO + (cb * FUSED_WAVES + wave) * span;  // first owned column
    if (x0 > g.nx - SDO) return false;                      // whole wavefront idle
    ft.col = x0 - SDO + lane;
    ft.colact = ft.col >= SDO && ft.col <= g.nx - SDO;
    ft.own = lane >= SDO && lane < 64 - SDO && ft.colact;
    ft.i0 = SDO + st * g.mrows;
    ft.i1 = min(ft.i0 + g.mrows - 1, g.nz - SDO);
    return true;
}

// x-derivatives from the neighbouring lanes' values (same row).
template <int SDO>
__device__ __forceinline__ float lane_dminus(float v, int lane, const float c[4])
{
    auto at = [&](int k) { return __shfl(v, lane + k, 64); };
    float r = c[0] * (v - at(-1));
    r = fmaf(c[1], at(1) - at(-2), r);
    if constexpr (SDO == 4) {
        r = fmaf(c[2], at(2) - at(-3), r);
        r = fmaf(c[3], at(3) - at(-4), r);
    }
    return r;
}

template <int SDO>
__device__ __forceinline__ float lane_dplus(float v, int lane, const float c[4])
{
    auto at = [&](int k) { return __shfl(v, lane + k, 64); };
    float r = c[0] * (at(1) - v);
    r = fmaf(c[1], at(2) - at(-1), r);
    if constexpr (SDO == 4) {
        r = fmaf(c[2], at(3) - at(-2), r);
        r = fmaf(c[3], at(4) - at(-3), r);
    }
    return r;
}

// Tile of a multi-stage launch (k_fwd_step, k_fwd_two, k_adj_fused): as
// tile_lane, with column blocks of SPAN = 64 - 2*H owned columns, H the halo
// lanes on each side (SDO per stage after the first); lane l of block cb
// sits on column cb*SPAN + l + SDO - H, and the halo lanes compute the early
// stages only.
template <int SDO, int SPB, int H = SDO>
__device__ __forceinline__ bool tile_fused(const DevGrid &g, int &s, int &j, int &i0, int &i1,
                                               bool &own, bool &colact)
{
    constexpr int SPAN = 64 - 2 * H;
    const int lane = (int)threadIdx.x % 64;
    int cb, st;
    if (SPB > 1 && g.nshot == 1) {
        const int ncq = (g.ncb + SPB - 1) / SPB;
        const int T = ncq * g.nstrips;
        const int per = (T + NUM_XCD - 1) / NUM_XCD;
        const int w = (int)blockIdx.x;
        const int logical = (w % NUM_XCD) * per + w / NUM_XCD;
        if (logical >= T) return false;
        cb = (logical % ncq) * SPB + wave_id();
        st = logical / ncq;
        s = 0;
        if (cb >= g.ncb) return false;
    } else {
        const int T = g.ncb * g.nstrips * g.nbatch;
        const int per = (T + NUM_XCD - 1) / NUM_XCD;
        const int w = (int)blockIdx.x;
        const int logical = (w % NUM_XCD) * per + w / NUM_XCD;
        if (logical >= T) return false;
        const int rest = logical / g.nbatch;
        cb = rest % g.ncb;
        st = rest / g.ncb;
        s = (logical % g.nbatch) * SPB + wave_id();
        if (s >= g.nshot) return false;
    }
    j = cb * SPAN + lane + SDO - H;
    colact = j >= SDO && j <= g.nx - SDO;
    own = lane >= H && lane < 64 - H && colact;
    i0 = SDO + st * g.mrows;
    i1 = min(i0 + g.mrows - 1, g.nz - SDO);
    return true;
}

// A lane's value at column offsets -SDO..SDO from the neighbouring lanes
// (wave shifts by one lane); the SDO edge lanes on each side get zeros from
// outside the wave, so only the owned lanes of a fused tile may use them.
template <int SDO>
struct LaneNbr {
    float s[2 * SDO + 1];
    __device__ __forceinline__ explicit LaneNbr(float c)
    {
        s[SDO] = c;
        float l = c, r = c;
#pragma unroll
        for (int k = 1; k <= SDO; ++k) {
            l = dpp_mov<DPP_WAVE_SHR1>(0.0f, l);  // lane i <- lane i-1
            r = dpp_mov<DPP_WAVE_SHL1>(0.0f, r);  // lane i <- lane i+1
            s[SDO - k] = l;
            s[SDO + k] = r;
        }
    }
    __device__ __forceinline__ float operator()(int k) const { return s[SDO + k]; }
};

// Forward step in one launch (src/PyFWI/elastic.cl:155-308): stage 1
// updates the velocities at row i, stage 2 the stresses at row q = i - SDO
// from stage 1's new velocities, held in register windows along z and taken
// from the neighbouring lanes (DPP) along x, so the new velocities are not
// read back from memory.  The state of step t is read from `in` and step t+1
// written to `out` (distinct buffers, per-shot strides in_stride and
// out_stride): the kept-state forward steps from one slot into the next.
// Per shot and cell: the state read and written once (40 B) and the model
// read once (12 B), 52 B against 72 B for k_vel + k_str.
//
// As k_vel and k_str, with `src` the step's source term is added to the
// values loaded at the source cell (vx, vz, txx, tzz; memory is not
// changed), and with `seis` the record sample of step t+1 is written at the
// receivers by the owning lane.  Every value is computed by the same
// arithmetic as in k_vel and k_str.  Stage 1 is forced to zero off the
// active cells, which stay zero in every state.
template <int SDO, int SPB>
__global__ __launch_bounds__(64 * SPB) void k_fwd_step(
    DevGrid g, const pyfwi_shot *shots, int t, const float *__restrict__ in, int64_t in_stride,
    float *__restrict__ out, int64_t out_stride, const float *__restrict__ bu,
    const float *__restrict__ lam, const float *__restrict__ mu, const float *__restrict__ dxp,
    const float *__restrict__ dzp, const float *__restrict__ src, float *seis)
{
    int s, j, i0, i1;
    bool own, colact;
    if (!tile_fused<SDO, SPB>(g, s, j, i0, i1, own, colact)) return;
    if (!fwd_active(shots, s, t)) return;
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    // bases 2*SDO rows above cell (0, 0): the byte offsets of every row the
    // kernel reads (from -2*SDO) are non-negative
    const int64_t org = g.origin - (int64_t)2 * SDO * g.pitch;
    const float *F = in + (int64_t)s * in_stride + org;
    float *O = out + (int64_t)s * out_stride + org;
    const CellRef vx{F}, vz{F + fs}, txx{F + 2 * fs}, tzz{F + 3 * fs}, txz{F + 4 * fs};
    const CellRef B{bu + org}, L{lam + org}, M{mu + org};
    // the step's source term on the values this lane loads (see k_vel)
    SrcTerm st = src_term(shots[s], t, src);
    if (st.z < i0 - 2 * SDO || st.z > i1 + 2 * SDO || st.x < j - SDO || st.x > j + SDO)
        st.z = -1;
    float s_txx[2 * SDO];  // txx term on the taps of d/dx- (offsets -SDO..SDO-1)
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) s_txx[k] = j + k - SDO == st.x ? st.v[2] : 0.0f;
    const bool src_col = j == st.x;
    const float s_vx = src_col ? st.v[0] : 0.0f, s_vz = src_col ? st.v[1] : 0.0f;
    const float s_tzz = src_col ? st.v[3] : 0.0f, s_txxc = src_col ? st.v[2] : 0.0f;
    RecLine rl = record_line(shots[s], t + 1, seis);
    const int rec_r = own && rl.z >= i0 && rl.z <= i1 ? rl.index(j) : -1;
    if (rec_r < 0) rl.z = -1;
    auto tzz_row = [&](uint32_t o, int row) {
        float v = tzz(o);
        if (row == st.z) v += s_tzz;
        return v;
    };
    auto dx_txz = [&](uint32_t o) { return xplus<SDO>([&](int k) { return txz.at(o, k); }, g.cx); };

    ZWin<SDO> wxz, wzz;    // old txz (d/dz-), old tzz (d/dz+) around stage 1's row
    Fifo<SDO> fxz;         // d/dx+ old txz, taken on the entering row
    Fifo<SDO> otxx;        // old txx (source included), delayed to stage 2's row
    ZWin<SDO> nvx{}, nvz{};  // new vx (d/dz+), new vz (d/dz-) around stage 2's row
    // first stage 1 row: i0 - SDO (the halo rows stage 2 needs above the strip)
    uint32_t cb = ((uint32_t)(i0 - SDO + 2 * SDO) * (uint32_t)g.pitch + (uint32_t)j) * 4u;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        const int row = i0 - SDO + k - SDO;
        wxz.v[k + 1] = txz(o);
        wzz.v[k + 1] = tzz_row(o, row);
        if (k >= SDO) fxz.v[k - SDO] = dx_txz(o);
        if (k < SDO) {
            // txx of stage 2's first rows, i0-2*SDO .. i0-SDO-1
            float v = txx(o);
            if (row == st.z) v += s_txxc;
            otxx.v[k] = v;
        }
    }
    const float ex = colact ? dxp[j] : 0.0f;
    const float dt = g.dt;
    const int nrow = i1 + SDO - (i0 - SDO) + 1;
    // rows in groups of MARCH_UNROLL, the inner loop fully unrolled (the lane
    // moves keep the compiler from unrolling a loop with a run-time trip
    // count); rows past the strip's end stay in the zero halo and store nothing
    for (int rr = 0; rr < nrow; rr += FWD_STEP_UNROLL) {
#pragma unroll
        for (int u = 0; u < FWD_STEP_UNROLL; ++u, cb += RB) {
            const int i = i0 - SDO + rr + u, q = i - SDO;
            const uint32_t n = cb + (uint32_t)SDO * RB, cq = cb - (uint32_t)SDO * RB;
            // loads of the row: stage 1's state and b, stage 2's model
            wxz.push(txz(n));
            wzz.push(tzz_row(n, i + SDO));
            const float dxp_txz = fxz.pop(dx_txz(n));
            Taps<SDO, 1> ttxx;
            ttxx.load(txx, cb);
            float ux = vx(cb), uz = vz(cb);
            const float bc = B(cb);
            const float l = L(cq), m = M(cq);
            if (i == st.z) {  // the source row
#pragma unroll
                for (int k = 0; k < 2 * SDO; ++k) ttxx.t[k] += s_txx[k];
                ux += s_vx;
                uz += s_vz;
            }
            const float a0 = otxx.pop(ttxx(0));  // old txx at row q
            // stage 1: velocities at row i (k_vel's arithmetic)
            const float dxm_txx = xminus<SDO>(ttxx, g.cx);
            const float ax = bc * (dxm_txx + wxz.dminus(g.cz));
            const float az = bc * (wzz.dplus(g.cz) + dxp_txz);
            const float damp1 = ex + dzp[i];
            ux = fmaf(dt, ax, ux);
            ux = fmaf(dt, -damp1 * ux, ux);
            uz = fmaf(dt, az, uz);
            uz = fmaf(dt, -damp1 * uz, uz);
            const bool act1 = colact && i >= SDO && i <= g.nz - SDO;
            ux = act1 ? ux : 0.0f;
            uz = act1 ? uz : 0.0f;
            nvx.push(ux);
            nvz.push(uz);
            // stage 2: stresses at row q (k_str's arithmetic), x-derivatives
            // of the new velocities from the neighbouring lanes
            const float exx = xplus<SDO>(LaneNbr<SDO>(nvx.centre()), g.cx);
            const float dxm_vz = xminus<SDO>(LaneNbr<SDO>(nvz.centre()), g.cx);
            const float ezz = nvz.dminus(g.cz);
            const float gxz = dxm_vz + nvx.dplus(g.cz);
            const float lm = l + 2.0f * m;
            const float sxx = fmaf(l, ezz, lm * exx);  // as in k_str
            const float szz = fmaf(l, exx, lm * ezz);
            const float sxz = m * gxz;
            const float damp2 = ex + dzp[q];
            float a = a0, bb = wzz.v[0], d = wxz.v[0];  // old stresses at row q
            a = fmaf(dt, sxx, a);
            bb = fmaf(dt, szz, bb);
            d = fmaf(dt, sxz, d);
            a = fmaf(dt, -damp2 * a, a);
            bb = fmaf(dt, -damp2 * bb, bb);
            d = fmaf(dt, -damp2 * d, d);
            // stores: the owned columns of the strip
            if (own && i >= i0 && i <= i1) {
                cell_ref(O, cb) = ux;
                cell_ref(O + fs, cb) = uz;
            }
            if (own && q >= i0 && q <= i1) {
                cell_ref(O + 2 * fs, cq) = a;
                cell_ref(O + 3 * fs, cq) = bb;
                cell_ref(O + 4 * fs, cq) = d;
                if (q == rl.z) {  // this lane's receiver
                    float *w = const_cast<float *>(rl.base) + rec_r;
                    w[0] = nvx.centre();
                    w[rl.fs] = nvz.centre();
                    w[2 * rl.fs] = a;
                    w[3 * rl.fs] = bb;
                    w[4 * rl.fs] = d;
                }
            }
        }
    }
}

// Two forward steps in one launch (time blocking, order 4): stages 1 and 2
// are k_fwd_step's for step t (velocities at row r, stresses at row
// r - SDO), stages 3 and 4 the same for step t+1 at rows r - 2*SDO and
// r - 3*SDO, each stage fed by the one before through register windows
// along z and lane moves along x.  State t-1 is read from `in`; the states
// after steps t and t+1 are written to `out0` and `out1` (per-shot stride
// out_stride), as two k_fwd_step launches would.  Per shot and cell and two
// steps: one state read (20 B), the model read once (12 B), two states
// written (40 B): 72 B against 104 B for two one-launch steps.  A wave owns
// the middle 64 - 6*SDO columns; each strip runs stage 1 on 3*SDO more rows
// above and below.  Step t+1's source term enters the values stages 3 and 4
// use, never the stored state t; every value is computed by the arithmetic
// of k_vel and k_str.  A shot whose last step is t stores nothing for t+1.
template <int SPB>
__global__ __launch_bounds__(64 * SPB) void k_fwd_two(
    DevGrid g, const pyfwi_shot *shots, int t, const float *__restrict__ in, int64_t in_stride,
    float *__restrict__ out0, float *__restrict__ out1, int64_t out_stride,
    const float *__restrict__ bu, const float *__restrict__ lam, const float *__restrict__ mu,
    const float *__restrict__ dxp, const float *__restrict__ dzp, const float *__restrict__ src,
    float *seis)
{
    constexpr int SDO = 2;
    int s, j, i0, i1;
    bool own, colact;
    if (!tile_fused<SDO, SPB, 3 * SDO>(g, s, j, i0, i1, own, colact)) return;
    if (!fwd_active(shots, s, t)) return;
    const bool second = fwd_active(shots, s, t + 1);  // wave-uniform
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    // bases 4*SDO rows above cell (0, 0): every row read (from -4*SDO) has a
    // non-negative byte offset
    const int64_t org = g.origin - (int64_t)4 * SDO * g.pitch;
    const float *F = in + (int64_t)s * in_stride + org;
    float *O0 = out0 + (int64_t)s * out_stride + org;
    float *O1 = out1 + (int64_t)s * out_stride + org;
    const CellRef vx{F}, vz{F + fs}, txx{F + 2 * fs}, tzz{F + 3 * fs}, txz{F + 4 * fs};
    const CellRef B{bu + org}, L{lam + org}, M{mu + org};
    // step t's source term on the loaded values (as k_fwd_step)
    SrcTerm s0 = src_term(shots[s], t, src);
    if (s0.z < i0 - 4 * SDO || s0.z > i1 + 4 * SDO || s0.x < j - SDO || s0.x > j + SDO)
        s0.z = -1;
    float s_txx[2 * SDO];  // txx term on the taps of d/dx- (offsets -SDO..SDO-1)
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) s_txx[k] = j + k - SDO == s0.x ? s0.v[2] : 0.0f;
    const bool src_col = j == s0.x;
    const float s_vx = src_col ? s0.v[0] : 0.0f, s_vz = src_col ? s0.v[1] : 0.0f;
    const float s_tzz = src_col ? s0.v[3] : 0.0f, s_txxc = src_col ? s0.v[2] : 0.0f;
    // step t+1's source term, added to stage 1 and 2's results at the source
    // cell before stages 3 and 4 use them
    SrcTerm s1 = second ? src_term(shots[s], t + 1, src) : SrcTerm{};
    if (s1.x != j) s1.z = -1;
    // record samples of steps t+1 and t+2, written by the owning lane
    RecLine r1 = record_line(shots[s], t + 1, seis);
    const int rec1 = own && r1.z >= i0 && r1.z <= i1 ? r1.index(j) : -1;
    if (rec1 < 0) r1.z = -1;
    RecLine r2 = second ? record_line(shots[s], t + 2, seis) : RecLine{};
    const int rec2 = own && r2.z >= i0 && r2.z <= i1 ? r2.index(j) : -1;
    if (rec2 < 0) r2.z = -1;
    auto tzz_row = [&](uint32_t o, int row) {
        float v = tzz(o);
        if (row == s0.z) v += s_tzz;
        return v;
    };
    auto dx_txz = [&](uint32_t o) { return xplus<SDO>([&](int k) { return txz.at(o, k); }, g.cx); };

    ZWin<SDO> wxz, wzz;        // state t-1: txz (d/dz-), tzz (d/dz+) around stage 1's row
    Fifo<SDO> fxz;             // d/dx+ txz, taken on the entering row
    Fifo<SDO> otxx;            // txx of state t-1 (source included), at stage 2's row
    ZWin<SDO> nvx{}, nvz{};    // stage 1's velocities around stage 2's row
    Fifo<2 * SDO> dvx{}, dvz{};  // stage 1's velocities (step t+1's source in), at stage 3's row
    ZWin<SDO> mtxz{}, mtzz{};  // stage 2's txz, tzz (source in) around stage 3's row
    Fifo<SDO> mtxx{};          // stage 2's txx (source in), at stage 3's row
    Fifo<SDO> mtxx2{};         // and at stage 4's row
    ZWin<SDO> wvx{}, wvz{};    // stage 3's velocities around stage 4's row
    // first stage 1 row: i0 - 3*SDO
    const int r0 = i0 - 3 * SDO;
    // (j < 0 on the halo lanes of the first column block: the offset is taken
    // in 64 bits, then is non-negative)
    uint32_t cb = (uint32_t)(((int64_t)(r0 + 4 * SDO) * g.pitch + j) * 4);
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        const int row = r0 + k - SDO;
        wxz.v[k + 1] = txz(o);
        wzz.v[k + 1] = tzz_row(o, row);
        if (k >= SDO) fxz.v[k - SDO] = dx_txz(o);
        if (k < SDO) {
            float v = txx(o);
            if (row == s0.z) v += s_txxc;
            otxx.v[k] = v;
        }
    }
    const float ex = colact ? dxp[j] : 0.0f;
    const float dt = g.dt;
    const int nrow = i1 + 3 * SDO - r0 + 1;
    for (int rr = 0; rr < nrow; rr += FWD_STEP_UNROLL) {
#pragma unroll
        for (int u = 0; u < FWD_STEP_UNROLL; ++u, cb += RB) {
            const int r = r0 + rr + u, q = r - SDO, p = r - 2 * SDO, o = r - 3 * SDO;
            // (wave-uniform) no row past the strip's end: the bottom strip's
            // next row would read below the padded fields
            if (r > i1 + 3 * SDO) break;
            const uint32_t n = cb + (uint32_t)SDO * RB;
            const uint32_t cq = cb - (uint32_t)SDO * RB, cp = cq - (uint32_t)SDO * RB,
                           co = cp - (uint32_t)SDO * RB;
            // stages 3 and 4 above the rows they feed (p < i0 - SDO, o < i0)
            // read the model and damping of row r instead: their results are
            // never used, and rows that far above the strip may lie outside
            // the padded fields
            const bool p_ok = p >= i0 - SDO, o_ok = o >= i0;
            const uint32_t cpl = p_ok ? cp : cb, col = o_ok ? co : cb;
            // loads of the row
            wxz.push(txz(n));
            wzz.push(tzz_row(n, r + SDO));
            const float dxp_txz = fxz.pop(dx_txz(n));
            Taps<SDO, 1> ttxx;
            ttxx.load(txx, cb);
            float ux = vx(cb), uz = vz(cb);
            const float b1 = B(cb), b3 = B(cpl);
            const float l2 = L(cq), m2 = M(cq), l4 = L(col), m4 = M(col);
            if (r == s0.z) {
#pragma unroll
                for (int k = 0; k < 2 * SDO; ++k) ttxx.t[k] += s_txx[k];
                ux += s_vx;
                uz += s_vz;
            }
            const float a0 = otxx.pop(ttxx(0));  // txx of state t-1 at row q
            // stage 1: step t's velocities at row r
            const float dxm_txx = xminus<SDO>(ttxx, g.cx);
            const float ax = b1 * (dxm_txx + wxz.dminus(g.cz));
            const float az = b1 * (wzz.dplus(g.cz) + dxp_txz);
            const float damp1 = ex + dzp[r];
            ux = fmaf(dt, ax, ux);
            ux = fmaf(dt, -damp1 * ux, ux);
            uz = fmaf(dt, az, uz);
            uz = fmaf(dt, -damp1 * uz, uz);
            const bool act1 = colact && r >= SDO && r <= g.nz - SDO;
            ux = act1 ? ux : 0.0f;
            uz = act1 ? uz : 0.0f;
            nvx.push(ux);
            nvz.push(uz);
            const bool src1_r = r == s1.z;
            const float vx3 = dvx.pop(src1_r ? ux + s1.v[0] : ux);  // at row p
            const float vz3 = dvz.pop(src1_r ? uz + s1.v[1] : uz);
            // stage 2: step t's stresses at row q
            const float exx = xplus<SDO>(LaneNbr<SDO>(nvx.centre()), g.cx);
            const float dxm_vz = xminus<SDO>(LaneNbr<SDO>(nvz.centre()), g.cx);
            const float ezz = nvz.dminus(g.cz);
            const float gxz = dxm_vz + nvx.dplus(g.cz);
            const float lm2 = l2 + 2.0f * m2;
            const float sxx = fmaf(l2, ezz, lm2 * exx);  // as in k_str
            const float szz = fmaf(l2, exx, lm2 * ezz);
            const float sxz = m2 * gxz;
            const float damp2 = ex + dzp[q];
            float a = a0, bb = wzz.v[0], d = wxz.v[0];
            a = fmaf(dt, sxx, a);
            bb = fmaf(dt, szz, bb);
            d = fmaf(dt, sxz, d);
            a = fmaf(dt, -damp2 * a, a);
            bb = fmaf(dt, -damp2 * bb, bb);
            d = fmaf(dt, -damp2 * d, d);
            const bool act2 = colact && q >= SDO && q <= g.nz - SDO;
            a = act2 ? a : 0.0f;
            bb = act2 ? bb : 0.0f;
            d = act2 ? d : 0.0f;
            const bool src1_q = q == s1.z;
            const float a_s = src1_q ? a + s1.v[2] : a, bb_s = src1_q ? bb + s1.v[3] : bb;
            mtxz.push(d);
            mtzz.push(bb_s);
            const float txx3 = mtxx.pop(a_s);     // at row p
            const float txx4 = mtxx2.pop(txx3);   // at row o
            // stage 3: step t+1's velocities at row p
            const float dxm_txx3 = xminus<SDO>(LaneNbr<SDO>(txx3), g.cx);
            const float dxp_txz3 = xplus<SDO>(LaneNbr<SDO>(mtxz.centre()), g.cx);
            const float ax3 = b3 * (dxm_txx3 + mtxz.dminus(g.cz));
            const float az3 = b3 * (mtzz.dplus(g.cz) + dxp_txz3);
            const float damp3 = ex + dzp[p_ok ? p : r];
            float ux3 = fmaf(dt, ax3, vx3);
            ux3 = fmaf(dt, -damp3 * ux3, ux3);
            float uz3 = fmaf(dt, az3, vz3);
            uz3 = fmaf(dt, -damp3 * uz3, uz3);
            const bool act3 = colact && p >= SDO && p <= g.nz - SDO;
            ux3 = act3 ? ux3 : 0.0f;
            uz3 = act3 ? uz3 : 0.0f;
            wvx.push(ux3);
            wvz.push(uz3);
            // stage 4: step t+1's stresses at row o
            const float exx4 = xplus<SDO>(LaneNbr<SDO>(wvx.centre()), g.cx);
            const float dxm_vz4 = xminus<SDO>(LaneNbr<SDO>(wvz.centre()), g.cx);
            const float ezz4 = wvz.dminus(g.cz);
            const float gxz4 = dxm_vz4 + wvx.dplus(g.cz);
            const float lm4 = l4 + 2.0f * m4;
            const float sxx4 = fmaf(l4, ezz4, lm4 * exx4);
            const float szz4 = fmaf(l4, exx4, lm4 * ezz4);
            const float sxz4 = m4 * gxz4;
            const float damp4 = ex + dzp[o_ok ? o : r];
            float a4 = txx4, bb4 = mtzz.v[0], d4 = mtxz.v[0];
            a4 = fmaf(dt, sxx4, a4);
            bb4 = fmaf(dt, szz4, bb4);
            d4 = fmaf(dt, sxz4, d4);
            a4 = fmaf(dt, -damp4 * a4, a4);
            bb4 = fmaf(dt, -damp4 * bb4, bb4);
            d4 = fmaf(dt, -damp4 * d4, d4);
            // stores: the owned columns of the strip
            if (own && r >= i0 && r <= i1) {
                cell_ref(O0, cb) = ux;
                cell_ref(O0 + fs, cb) = uz;
            }
            if (own && q >= i0 && q <= i1) {
                cell_ref(O0 + 2 * fs, cq) = a;
                cell_ref(O0 + 3 * fs, cq) = bb;
                cell_ref(O0 + 4 * fs, cq) = d;
                if (q == r1.z) {
                    float *w = const_cast<float *>(r1.base) + rec1;
                    w[0] = nvx.centre();
                    w[r1.fs] = nvz.centre();
                    w[2 * r1.fs] = a;
                    w[3 * r1.fs] = bb;
                    w[4 * r1.fs] = d;
                }
            }
            if (second && own && p >= i0 && p <= i1) {
                cell_ref(O1, cp) = ux3;
                cell_ref(O1 + fs, cp) = uz3;
            }
            if (second && own && o >= i0 && o <= i1) {
                cell_ref(O1 + 2 * fs, co) = a4;
                cell_ref(O1 + 3 * fs, co) = bb4;
                cell_ref(O1 + 4 * fs, co) = d4;
                if (o == r2.z) {
                    float *w = const_cast<float *>(r2.base) + rec2;
                    w[0] = wvx.centre();
                    w[r2.fs] = wvz.centre();
                    w[2 * r2.fs] = a4;
                    w[3 * r2.fs] = bb4;
                    w[4 * r2.fs] = d4;
                }
            }
        }
    }
}

// Time-reversed step (coeff = -1, wave_propagation.py:780-812): stage 1
// stresses, stage 2 velocities, damping before each update.
template <int SDO>
__global__ __launch_bounds__(64 * FUSED_WAVES) void k_fused_rev(
    DevGrid g, const pyfwi_shot *shots, int t, const float *in, float *out,
    const float *__restrict__ bu, const float *__restrict__ lam, const float *__restrict__ mu,
    const float *__restrict__ dxp, const float *__restrict__ dzp)
{
    FusedTile<SDO> ft;
    if (!fused_tile<SDO>(g, ft)) return;
    if (!bwd_active(shots, ft.s, t)) return;
    const int64_t fs = g.fstride;
    const int P = g.pitch, lane = (int)threadIdx.x % 64, col = ft.col;
    const float *I = in + (int64_t)ft.s * 5 * fs + g.origin;
    float *O = out + (int64_t)ft.s * 5 * fs + g.origin;
    const float *vx = I, *vz = I + fs, *txx = I + 2 * fs, *tzz = I + 3 * fs, *txz = I + 4 * fs;
    const float *b = bu + g.origin, *L = lam + g.origin, *M = mu + g.origin;
    const bool ld = ft.colact;
    ZWin<SDO> ovz{}, ovx{};      // old vz (d/dz-), old vx (d/dz+): rows r-SDO..r+SDO
    ZWin<SDO> ntxx{}, ntzz{}, ntxz{};  // new stresses: rows q-SDO..q+SDO
    const int rs = ft.i0 - SDO;
    const int64_t c0 = (int64_t)rs * P + col;
    int64_t c = c0;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const int64_t o = c + (int64_t)(k - SDO) * P;
        ovz.v[k + 1] = ld ? vz[o] : 0.0f;
        ovx.v[k + 1] = ld ? vx[o] : 0.0f;
    }
    const float ddx = ld ? dxp[col] : 0.0f;
    const float dt = g.dt;
    // rows in groups of MARCH_UNROLL, the inner loop fully unrolled: the cross-lane
    // shuffles stop the compiler from unrolling a loop with a run-time trip count.
    // Rows past the strip end (at most MARCH_UNROLL - 1) stay inside the zero halo
    // and write nothing.
    const int nrow = ft.i1 + SDO - rs + 1;
    for (int rr = 0; rr < nrow; rr += MARCH_UNROLL) {
#pragma unroll
    for (int u = 0; u < MARCH_UNROLL; ++u) {
        const int r = rs + rr + u;
        c = c0 + (int64_t)(rr + u) * P;
        const int64_t n = c + (int64_t)SDO * P;
        ovz.push(ld ? vz[n] : 0.0f);
        ovx.push(ld ? vx[n] : 0.0f);
        float a = 0.0f, bb = 0.0f, d = 0.0f;
        const bool act1 = ld && r >= SDO && r <= g.nz - SDO;
        if (act1) {
            const float exx = dplus<SDO>(vx + c, 1, g.cx);
            const float ezz = ovz.dminus(g.cz);
            const float gxz = dminus<SDO>(vz + c, 1, g.cx) + ovx.dplus(g.cz);
            const float l = L[c], m = M[c];
            const float lm = l + 2.0f * m;
            const float damp = ddx + dzp[r];
            a = txx[c];
            bb = tzz[c];
            d = txz[c];
            a = fmaf(-dt, damp * a, a);
            bb = fmaf(-dt, damp * bb, bb);
            d = fmaf(-dt, damp * d, d);
            a = fmaf(-dt, lm * exx + l * ezz, a);
            bb = fmaf(-dt, lm * ezz + l * exx, bb);
            d = fmaf(-dt, m * gxz, d);
            if (ft.own && r >= ft.i0 && r <= ft.i1) {
                O[2 * fs + c] = a;
                O[3 * fs + c] = bb;
                O[4 * fs + c] = d;
            }
        }
        ntxx.push(a);
        ntzz.push(bb);
        ntxz.push(d);
        const int q = r - SDO;
        const float dxm_txx = lane_dminus<SDO>(ntxx.centre(), lane, g.cx);
        const float dxp_txz = lane_dplus<SDO>(ntxz.centre(), lane, g.cx);
        if (q < ft.i0 || q > ft.i1 || !ft.own) continue;
        const int64_t cq = c - (int64_t)SDO * P;
        const float ax = b[cq] * (dxm_txx + ntxz.dminus(g.cz));
        const float az = b[cq] * (ntzz.dplus(g.cz) + dxp_txz);
        const float damp = ddx + dzp[q];
        float ux = ovx.v[0], uz = ovz.v[0];
        ux = fmaf(-dt, damp * ux, ux);
        ux = fmaf(-dt, ax, ux);
        uz = fmaf(-dt, damp * uz, uz);
        uz = fmaf(-dt, az, uz);
        O[cq] = ux;
        O[fs + cq] = uz;
    }
    }
}

// Offset from cell (0, 0) of a padded field of receiver-table entry `e`
// (row*nx + column of the unpadded grid).
__device__ __forceinline__ int64_t table_cell(const DevGrid &g, int32_t e)
{
    const int row = e / g.nx;
    return (int64_t)row * g.pitch + (e - row * g.nx);
}

// Field a record component samples: the crosswell gather fills the vz
// record from vx (src/PyFWI/elastic_crosswell.cl:27).
__device__ __forceinline__ int recorded_field(const pyfwi_shot &sh, int f)
{
    return (f == 1 && (sh.rec_flags & PYFWI_REC_VZ_FROM_VX)) ? 0 : f;
}

// Record row k of one shot from the state F (cell (0, 0) of its first field;
// nullptr: the zero state before the first step), the block's threads
// striding over the receivers.
__device__ __forceinline__ void record_row(const DevGrid &g, const pyfwi_shot &sh, int k,
                                           const float *F, float *seis,
                                           const int32_t *__restrict__ rtab)
{
    const int64_t fs = g.fstride;
    float *rec = seis + sh.rec_offset;
    const int32_t *cells = sh.rec_layout == PYFWI_REC_TABLE ? rtab + sh.rec_table : nullptr;
    const int64_t c0 = (int64_t)sh.rec_z * g.pitch + sh.rec_x0;
    for (int r = threadIdx.x; r < sh.n_rec; r += blockDim.x) {
        int64_t c = c0 + (int64_t)r * sh.rec_dx;
        bool live = F != nullptr;
        if (cells) {
            const int32_t e = cells[r];
            live = live && e >= 0;
            c = e >= 0 ? table_cell(g, e) : 0;
        }
#pragma unroll
        for (int f = 0; f < 5; ++f)
            rec[((int64_t)f * sh.n_samples + k) * sh.rec_stride + r] =
                live ? F[recorded_field(sh, f) * fs + c] : 0.0f;
    }
}

// Receiver gather then source injection of step t for every shot, one block
// per shot (src/PyFWI/elastic_surface.cl:7-41; the well gathers of
// elastic_crosswell.cl:1-33 and elastic.cl:43-98 through the receiver
// table).  Record row k = t/dt_scale holds the state after t steps; the
// source sample k is added to vx, vz, txx, tzz at the source cell afterwards,
// unless src == nullptr (gather only, the stencil kernels then apply the
// source term).  `fields`: per-shot stride fields_stride.
__global__ void k_inject_record(DevGrid g, const pyfwi_shot *shots, int t,
                                float *fields, int64_t fields_stride,
                                const float *__restrict__ src, float *seis,
                                const int32_t *__restrict__ rtab)
{
    const int s = blockIdx.x;
    const pyfwi_shot sh = shots[s];
    if (t >= sh.nt || (t % sh.dt_scale) != 0) return;
    const int k = t / sh.dt_scale;
    const int64_t fs = g.fstride;
    float *F = fields + (int64_t)s * fields_stride + g.origin;
    if (k < sh.n_samples) record_row(g, sh, k, F, seis, rtab);
    if (!src) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float *sv = src + sh.src_offset + (int64_t)k * 5;
        const int64_t c = (int64_t)sh.src_z * g.pitch + sh.src_x;
        F[c] += sv[0];
        F[fs + c] += sv[1];
        F[2 * fs + c] += sv[2];
        F[3 * fs + c] += sv[3];
    }
}

// Records of steps t0 .. t0+gridDim.x-1 gathered after a forward that kept
// every state, one block per (step, shot): the gather of step t reads the
// state step t-1 left in its slot (slot t-1 of shot s at
// kept + s*shot_stride + (t-1)*slot_stride), the state k_inject_record reads
// before step t; kept slots never hold a source term.
__global__ void k_record_kept(DevGrid g, const pyfwi_shot *shots, int t0, const float *kept,
                              int64_t shot_stride, int64_t slot_stride, float *seis,
                              const int32_t *__restrict__ rtab)
{
    const int t = t0 + (int)blockIdx.x;
    const int s = blockIdx.y;
    const pyfwi_shot sh = shots[s];
    if (t >= sh.nt || (t % sh.dt_scale) != 0) return;
    const int k = t / sh.dt_scale;
    if (k >= sh.n_samples) return;
    const float *F = t > 0 ? kept + (int64_t)s * shot_stride + (int64_t)(t - 1) * slot_stride +
                                 g.origin
                           : nullptr;
    record_row(g, sh, k, F, seis, rtab);
}

// ---------------------------------------------------- adjoint, pyfwi mode

// Residual injection in memory at backward step t (Adj_injSrc,
// elastic_surface.cl:48-72, elastic_crosswell.cl:35-58, elastic.cl:316-363),
// components f0..f1-1, component f into adjoint field f; with `first_only`,
// only for the shots whose first backward step t is (t == nt - 1), the
// adjoint velocities of later steps being written with their residual by the
// velocity update of the step before.  Table receivers may share a cell (a
// well's top receiver on the surface line): their residuals are summed.
__global__ void k_adj_inject(DevGrid g, const pyfwi_shot *shots, int t,
                             float *adj, const float *__restrict__ res, int f0, int f1,
                             int first_only, const int32_t *__restrict__ rtab)
{
    const int s = blockIdx.x;
    const pyfwi_shot sh = shots[s];
    if (first_only && t != sh.nt - 1) return;
    const RecLine rq = residual_line(sh, t, res);
    if (!rq.base) return;
    const int64_t fs = g.fstride;
    float *A = adj + (int64_t)s * 5 * fs + g.origin;
    if (sh.rec_layout == PYFWI_REC_TABLE) {
        const int32_t *cells = rtab + sh.rec_table + sh.n_rec;
        for (int r = threadIdx.x; r < sh.n_rec; r += blockDim.x) {
            const int32_t e = cells[r];
            if (e < 0) continue;
            const int64_t c = table_cell(g, e);
            for (int f = f0; f < f1; ++f) atomicAdd(&A[f * fs + c], rq.base[f * rq.fs + r]);
        }
        return;
    }
    const int64_t c0 = (int64_t)sh.rec_z * g.pitch + sh.rec_x0;
    for (int r = threadIdx.x; r < sh.n_rec; r += blockDim.x) {
        const int64_t c = c0 + (int64_t)r * sh.rec_dx;
        for (int f = f0; f < f1; ++f) A[f * fs + c] += rq.base[f * rq.fs + r];
    }
}

// The imaging condition of a step (elastic.cl:470-526) reads the adjoint
// fields only at the cell itself, so its terms go where those fields are
// made: the lambda and mu parts, which need the new adjoint stresses and the
// forward velocities' derivatives, in the adjoint stress update; the rho
// part, which needs the new adjoint velocities and the forward stresses'
// derivatives, in the adjoint velocity update.  Each kernel then reads about
// the same number of arrays.  The shear term uses the z-derivative of vz as
// the reference does (elastic.cl:485-486).
//
// The shots of a block (one wave each, same cells) add their terms up before
// they reach memory: each wave stages a shot's terms for ADJ_GROUP_ROWS rows
// in LDS, and after a barrier the live waves share those rows, each adding
// the shots' terms in shot order onto the group's accumulators.  The
// accumulators then move once per group instead of once per shot.  gacc:
// [adj_acc_sets(sdo, nshots)][6][fstride] = gmu, gmu_pre, glam, glam_pre, grho,
// grho_pre; interior cells only.

// The shots of the calling wave's group that take part in step t (waves of
// the others have left the kernel), and the wave's rank among them.
struct GroupLive {
    uint32_t mask = 0;
    int count = 0, rank = 0;
};

// FROM0: the exact adjoint's range of steps, 0 <= t < nt.
template <int SPB, bool FROM0 = false>
__device__ __forceinline__ GroupLive group_live(const DevGrid &g, const pyfwi_shot *shots, int s,
                                                int t)
{
    const int wave = wave_id(), s0 = s - wave;
    GroupLive r;
#pragma unroll
    for (int w = 0; w < SPB; ++w) {
        const bool on = FROM0 ? t < shots[s0 + w < g.nshot ? s0 + w : s].nt
                              : s0 + w < g.nshot && bwd_active(shots, s0 + w, t);
        if (s0 + w < g.nshot && on) r.mask |= 1u << w;
    }
    r.count = __popc(r.mask);
    r.rank = __popc(r.mask & ((1u << wave) - 1u));
    return r;
}

// Adds the staged terms of rows ic .. ic+ADJ_GROUP_ROWS-1 (part: [row][shot]
// [term][lane]) onto the group accumulators G + slot[v]*fs; rows are dealt
// out to the live waves by rank.  `c0`: byte offset of (row ic, the lane's
// column).  Called by every live wave after the barrier that follows the
// staging.
template <int NV, int SPB>
__device__ __forceinline__ void add_group_terms(const float *part, const GroupLive &gl, int ic,
                                                int i1, bool img, int lo, int hi, float *G,
                                                const int *slot, int64_t fs, uint32_t c0,
                                                uint32_t RB)
{
    const int lane = (int)threadIdx.x % 64;
    for (int r = gl.rank; r < ADJ_GROUP_ROWS; r += gl.count) {
        const int i = ic + r;
        if (i > i1) break;
        if (!img || i < lo || i > hi) continue;
        const uint32_t o = c0 + (uint32_t)r * RB;
        float cur[NV], acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            cur[v] = cell_ref(G + slot[v] * fs, o);
            acc[v] = 0.0f;
        }
#pragma unroll
        for (int w = 0; w < SPB; ++w) {
            if (!(gl.mask >> w & 1u)) continue;
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += part[((r * SPB + w) * NV + v) * 64 + lane];
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) cell_ref(G + slot[v] * fs, o) = cur[v] + acc[v];
    }
}

// Adjoint stresses from adjoint velocities (src/PyFWI/elastic.cl:372-411) and
// the lambda / mu imaging terms.
template <int SDO, bool PRECOND, bool ALONE, int SPB>
__global__ __launch_bounds__(64 * SPB) void k_adj_str(DevGrid g, const pyfwi_shot *shots, int t,
                                                     const float *fields, int64_t fields_stride,
                                                     float *adj, const float *__restrict__ bu,
                                                     const float *__restrict__ dxp,
                                                     const float *__restrict__ dzp, float *gacc,
                                                     const float *__restrict__ res)
{
    constexpr int NV = PRECOND ? 4 : 2;  // staged terms: mu, lam (, mu_pre, lam_pre)
    __shared__ float part[2][ADJ_GROUP_ROWS * SPB * NV * 64];
    int s, j, i0, i1;
    if (!tile<SDO, SPB>(g, s, j, i0, i1)) return;
    if (!bwd_active(shots, s, t)) return;
    GroupLive gl;  // the shot group (one-shot launches have none)
    if constexpr (!ALONE) gl = group_live<SPB>(g, shots, s, t);
    const int wave = wave_id(), lane = (int)threadIdx.x % 64;
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    float *A = adj + (int64_t)s * 5 * fs + g.origin;
    const CellRef avx{A}, avz{A + fs}, B{bu + g.origin};
    // the step's residual on the adjoint stresses of this thread's receiver
    // (zero elsewhere), added before the update
    const RecLine rq = residual_line(shots[s], t, res);
    const int rr = rq.z >= i0 && rq.z <= i1 ? rq.index(j) : -1;
    const int rrow = rr >= 0 ? rq.z : -1;
    const float r_txx = rr >= 0 ? rq.base[2 * rq.fs + rr] : 0.0f;
    const float r_tzz = rr >= 0 ? rq.base[3 * rq.fs + rr] : 0.0f;
    const float r_txz = rr >= 0 ? rq.base[4 * rq.fs + rr] : 0.0f;
    const int lo = g.npml + SDO, hi = g.nz - SDO - g.npml, hix = g.nx - SDO - g.npml;
    const bool img = j >= lo && j <= hix && i1 >= lo && i0 <= hi;
    const float *F = fields + (int64_t)s * fields_stride + g.origin;
    const CellRef vx{F}, vz{F + fs}, txx{F + 2 * fs}, tzz{F + 3 * fs}, txz{F + 4 * fs};
    float *G = gacc + (int64_t)(s / SPB) * 6 * fs + g.origin;
    constexpr int slots[4] = {0, 2, 1, 3};  // accumulator of each staged term
    ZWin<SDO> wbz, wbx;  // b*avz for d/dz-, b*avx for d/dz+
    Fifo<SDO> fxp, fxm;  // d/dx+ (b*avx), d/dx- (b*avz)
    ZWin<SDO> fvz;       // forward vz (d/dz+ and d/dz-)
    Fifo<SDO> fdvz;      // d/dx- of forward vz
    auto dx_bvx = [&](uint32_t o) {
        return xplus<SDO>([&](int k) { return avx.at(o, k) * B.at(o, k); }, g.cx); };
    auto dx_bvz = [&](uint32_t o) {
        return xminus<SDO>([&](int k) { return avz.at(o, k) * B.at(o, k); }, g.cx); };
    auto x_vz = [&](uint32_t o) { return xminus<SDO>([&](int k) { return vz.at(o, k); }, g.cx); };
    uint32_t cb = ((uint32_t)i0 * (uint32_t)g.pitch + (uint32_t)j) * 4u;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        wbz.v[k + 1] = avz(o) * B(o);
        wbx.v[k + 1] = avx(o) * B(o);
        if (k >= SDO) {
            fxp.v[k - SDO] = dx_bvx(o);
            fxm.v[k - SDO] = dx_bvz(o);
        }
        if (img) {
            fvz.v[k + 1] = vz(o);
            if (k >= SDO) fdvz.v[k - SDO] = x_vz(o);
        }
    }
    const float ex = dxp[j];
    const float dt = g.dt;
    // ALONE (launches of one shot): the shot adds its terms onto the
    // accumulators itself, with no staging and no barrier; the sums are the
    // same as a group of one gets through the staging.
    int buf = 0;
    for (int ic = i0; ic <= i1; ic += ADJ_GROUP_ROWS, buf ^= 1) {
        const uint32_t c0 = cb;
        float *mine = &part[buf][wave * NV * 64 + lane];
        // Every load of a row is issued before the row's stores: the compiler
        // does not move a load across a store to memory it cannot prove
        // distinct, and a load placed after one waits out a full memory latency.
#pragma unroll MARCH_UNROLL
        for (int r = 0; r < ADJ_GROUP_ROWS; ++r, cb += RB) {
            const int i = ic + r;
            if (i > i1) break;
            const uint32_t n = cb + (uint32_t)SDO * RB;
            wbz.push(avz(n) * B(n));
            wbx.push(avx(n) * B(n));
            const float dxpb = fxp.pop(dx_bvx(n));
            const float dxmb = fxm.pop(dx_bvz(n));
            const float ez = dzp[i];
            float *pa = &cell_ref(A + 2 * fs, cb), *pb = &cell_ref(A + 3 * fs, cb),
                  *pd = &cell_ref(A + 4 * fs, cb);
            const float a0 = *pa, b0 = *pb, d0 = *pd;
            const bool row_img = img && i >= lo && i <= hi;
            float vz_n = 0.0f;
            float sxx = 0.0f, szz = 0.0f, sxz = 0.0f;
            Taps<SDO, 1> tvz;  // forward vz on the entering row (d/dx-)
            Taps<SDO, 0> tvx;  // forward vx on the centre row (d/dx+)
            if (img) {
                vz_n = vz(n);
                tvz.load(vz, n);
            }
            float gc[NV] = {};  // the row's accumulators (ALONE)
            if (row_img) {
                tvx.load(vx, cb);
                if constexpr (ALONE) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) gc[v] = cell_ref(G + slots[v] * fs, cb);
                }
                if constexpr (PRECOND) {
                    sxx = txx(cb);
                    szz = tzz(cb);
                    sxz = txz(cb);
                }
            }
            // (lanes outside the imaging region push values they never use)
            fvz.push(vz_n);
            const float dxm_vz = fdvz.pop(xminus<SDO>(tvz, g.cx));
            const float dxp_vx = xplus<SDO>(tvx, g.cx);
            const float dzmb = wbz.dminus(g.cz);
            const float dzpb = wbx.dplus(g.cz);
            const bool at_rec = i == rrow;
            const float a = at_rec ? a0 + r_txx : a0;
            const float bb = at_rec ? b0 + r_tzz : b0;
            const float d = at_rec ? d0 + r_txz : d0;
            const float aa = fmaf(-dt, ex * a + dxpb, a);
            const float ab = fmaf(-dt, ez * bb + dzmb, bb);
            const float ad = fmaf(-dt, (ex + ez) * d + (dzpb + dxmb), d);
            *pa = aa;
            *pb = ab;
            *pd = ad;
            if (!row_img) continue;
            const float dzp_vz = fvz.dplus(g.cz);
            const float dzm_vz = fvz.dminus(g.cz);
            float term[NV];
            term[0] = 2.0f * dt * (aa * dxp_vx + ab * dzm_vz) + dt * ad * (dzp_vz + dxm_vz);
            term[1] = dt * (aa + ab) * (dxp_vx + dzm_vz);
            if constexpr (PRECOND) {
                const float dt2 = dt * dt;
                term[2] = 4.0f * dt2 * (sxx * sxx * dxp_vx * dxp_vx + szz * szz * dzm_vz * dzm_vz)
                          + dt2 * sxz * sxz * (dzp_vz * dzp_vz + dxm_vz * dxm_vz);
                term[3] = dt2 * (sxx * sxx + szz * szz) * (dxp_vx * dxp_vx + dzm_vz * dzm_vz);
            }
            float *q = mine + r * SPB * NV * 64;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if constexpr (ALONE) cell_ref(G + slots[v] * fs, cb) = gc[v] + term[v];
                else q[v * 64] = term[v];
            }
        }
        if constexpr (!ALONE) {
            __syncthreads();
            add_group_terms<NV, SPB>(part[buf], gl, ic, i1, img, lo, hi, G, slots, fs, c0, RB);
        }
    }
}

// Adjoint velocities (src/PyFWI/elastic.cl:419-468) and the rho imaging term.
template <int SDO, bool PRECOND, bool ALONE, int SPB>
__global__ __launch_bounds__(64 * SPB) void k_adj_vel_grad(
    DevGrid g, const pyfwi_shot *shots, int t, const float *fields, int64_t fields_stride,
    float *adj, const float *__restrict__ lam, const float *__restrict__ mu,
    const float *__restrict__ bu, const float *__restrict__ dxp,
    const float *__restrict__ dzp, float *gacc, const float *__restrict__ res)
{
    constexpr int NV = PRECOND ? 2 : 1;  // staged terms: rho (, rho_pre)
    __shared__ float part[2][ADJ_GROUP_ROWS * SPB * NV * 64];
    // order 8 takes its x-taps from lane shifts: every lane of a wave stays,
    // and only those on the grid store.  (In k_adj_str, with three such
    // derivatives a row, the same change was 18 % slower at order 8; it
    // keeps its tap loads.)
    constexpr bool LT = SDO == 4;
    int s, j, i0, i1;
    bool lane_ok = true;
    if constexpr (LT) {
        if (!tile_lane<SDO, SPB>(g, s, j, i0, i1, lane_ok)) return;
    } else if (!tile<SDO, SPB>(g, s, j, i0, i1)) {
        return;
    }
    if (!bwd_active(shots, s, t)) return;
    GroupLive gl;  // the shot group (one-shot launches have none)
    if constexpr (!ALONE) gl = group_live<SPB>(g, shots, s, t);
    const int wave = wave_id(), lane = (int)threadIdx.x % 64;
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    float *A = adj + (int64_t)s * 5 * fs + g.origin;
    const CellRef atxx{A + 2 * fs}, atzz{A + 3 * fs}, atxz{A + 4 * fs};
    // the residual of the next backward step (t-1) on the adjoint velocities
    // of this thread's receiver, added to what is written (this step's
    // imaging uses the velocities without it)
    const RecLine rq = residual_line(shots[s], t - 1, res);
    const int rr = lane_ok && rq.z >= i0 && rq.z <= i1 ? rq.index(j) : -1;
    const int rrow = rr >= 0 ? rq.z : -1;
    const float r_vx = rr >= 0 ? rq.base[rr] : 0.0f;
    const float r_vz = rr >= 0 ? rq.base[rq.fs + rr] : 0.0f;
    const CellRef L{lam + g.origin}, M{mu + g.origin}, B{bu + g.origin};
    const int lo = g.npml + SDO, hi = g.nz - SDO - g.npml, hix = g.nx - SDO - g.npml;
    const bool img = j >= lo && j <= hix && i1 >= lo && i0 <= hi;
    // forward-state loads: with lane shifts, on every lane of a wave that has
    // an imaging lane
    const bool imgl = LT ? __ballot(img) != 0 : img;
    const uint32_t hoff = halo_offset<SDO>(lane);
    const float *F = fields + (int64_t)s * fields_stride + g.origin;
    const CellRef vx{F}, vz{F + fs}, txx{F + 2 * fs}, tzz{F + 3 * fs}, txz{F + 4 * fs};
    float *G = gacc + (int64_t)(s / SPB) * 6 * fs + g.origin;
    constexpr int slots[2] = {4, 5};  // accumulator of each staged term

    ZWin<SDO> wa3, wb1, wb2;  // mu*atxz (d/dz-), (lam+2mu)*atzz, lam*atxx (d/dz+)
    ZWin<SDO> ftzz, ftxz;     // forward tzz (d/dz+), txz (d/dz-)
    // x-derivatives of the windowed fields, delayed to the centre row
    Fifo<SDO> fa1, fa2, fb3, fdtxz;
    auto q_a3 = [&](uint32_t o) { return M(o) * atxz(o); };
    auto q_b1 = [&](uint32_t o) { return (L(o) + 2.0f * M(o)) * atzz(o); };
    auto q_b2 = [&](uint32_t o) { return L(o) * atxx(o); };
    auto x_a1 = [&](uint32_t o) {
        if constexpr (LT) {
            auto q = [&](uint32_t p) { return (L(p) + 2.0f * M(p)) * atxx(p); };
            return xminus<SDO>(LaneRow<SDO>(q(o), q(o + hoff)), g.cx);
        } else {
            return xminus<SDO>([&](int k) {
                return (L.at(o, k) + 2.0f * M.at(o, k)) * atxx.at(o, k); }, g.cx);
        }
    };
    auto x_a2 = [&](uint32_t o) {
        if constexpr (LT)
            return xminus<SDO>(LaneRow<SDO>(L(o) * atzz(o), L(o + hoff) * atzz(o + hoff)), g.cx);
        else
            return xminus<SDO>([&](int k) { return L.at(o, k) * atzz.at(o, k); }, g.cx);
    };
    auto x_b3 = [&](uint32_t o) {
        if constexpr (LT)
            return xplus<SDO>(LaneRow<SDO>(M(o) * atxz(o), M(o + hoff) * atxz(o + hoff)), g.cx);
        else
            return xplus<SDO>([&](int k) { return M.at(o, k) * atxz.at(o, k); }, g.cx);
    };
    auto x_txz = [&](uint32_t o) {
        if constexpr (LT)
            return xplus<SDO>(LaneRow<SDO>(txz(o), txz(o + hoff)), g.cx);
        else
            return xplus<SDO>([&](int k) { return txz.at(o, k); }, g.cx);
    };
    uint32_t cb = ((uint32_t)i0 * (uint32_t)g.pitch + (uint32_t)j) * 4u;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        wa3.v[k + 1] = q_a3(o);
        wb1.v[k + 1] = q_b1(o);
        wb2.v[k + 1] = q_b2(o);
        if (k >= SDO) {
            const int q = k - SDO;
            fa1.v[q] = x_a1(o);
            fa2.v[q] = x_a2(o);
            fb3.v[q] = x_b3(o);
        }
        if (imgl) {
            ftzz.v[k + 1] = tzz(o);
            ftxz.v[k + 1] = txz(o);
            if (k >= SDO) fdtxz.v[k - SDO] = x_txz(o);
        }
    }
    const float ex = lane_ok ? dxp[j] : 0.0f;
    const float dt = g.dt;
    // ALONE: the shot adds its terms itself (see k_adj_str)
    int buf = 0;
    for (int ic = i0; ic <= i1; ic += ADJ_GROUP_ROWS, buf ^= 1) {
        const uint32_t c0 = cb;
        float *mine = &part[buf][wave * NV * 64 + lane];
        // every load of a row before the row's stores (see k_adj_str)
#pragma unroll ADJ_VEL_UNROLL
        for (int r = 0; r < ADJ_GROUP_ROWS; ++r, cb += RB) {
            const int i = ic + r;
            if (i > i1) break;
            const uint32_t n = cb + (uint32_t)SDO * RB;
            wa3.push(q_a3(n));
            wb1.push(q_b1(n));
            wb2.push(q_b2(n));
            const float a1 = fa1.pop(x_a1(n)), a2 = fa2.pop(x_a2(n));
            const float b3 = fb3.pop(x_b3(n));
            const float ez = dzp[i];
            float *px = &cell_ref(A, cb), *pz = &cell_ref(A + fs, cb);
            float ux = *px, uz = *pz;
            const bool row_img = img && i >= lo && i <= hi;
            const bool row_imgl = imgl && i >= lo && i <= hi;
            float tzz_n = 0.0f, txz_n = 0.0f, txz_h = 0.0f, txx_c = 0.0f, txx_h = 0.0f, bb = 0.0f;
            float wx = 0.0f, wz = 0.0f;
            Taps<SDO, 0> ttxz;  // forward txz on the entering row (d/dx+)
            Taps<SDO, 1> ttxx;  // forward txx on the centre row (d/dx-)
            if (imgl) {
                tzz_n = tzz(n);
                txz_n = txz(n);
                if constexpr (LT) txz_h = txz(n + hoff);
                else ttxz.load(txz, n);
            }
            float gc[NV] = {};  // the row's accumulators (ALONE)
            if (row_imgl) {
                if constexpr (LT) {
                    txx_c = txx(cb);
                    txx_h = txx(cb + hoff);
                } else {
                    ttxx.load(txx, cb);
                }
                bb = B(cb);
                if constexpr (ALONE) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) gc[v] = cell_ref(G + slots[v] * fs, cb);
                }
                if constexpr (PRECOND) {
                    wx = vx(cb);
                    wz = vz(cb);
                }
            }
            // (lanes outside the imaging region push values they never use)
            ftzz.push(tzz_n);
            ftxz.push(txz_n);
            float dxp_txz_n, dxm_txx;
            if constexpr (LT) {
                dxp_txz_n = xplus<SDO>(LaneRow<SDO>(txz_n, txz_h), g.cx);
                dxm_txx = xminus<SDO>(LaneRow<SDO>(txx_c, txx_h), g.cx);
            } else {
                dxp_txz_n = xplus<SDO>(ttxz, g.cx);
                dxm_txx = xminus<SDO>(ttxx, g.cx);
            }
            const float dxp_txz = fdtxz.pop(dxp_txz_n);
            const float a3 = wa3.dminus(g.cz);
            const float b1 = wb1.dplus(g.cz);
            const float b2 = wb2.dplus(g.cz);
            ux = fmaf(-dt, (2.0f * ex + ez) * ux + (a1 + a2 + a3), ux);
            uz = fmaf(-dt, (ex + 2.0f * ez) * uz + (b2 + b1 + b3), uz);
            const bool at_rec = i == rrow;
            if (lane_ok) {
                *px = at_rec ? ux + r_vx : ux;
                *pz = at_rec ? uz + r_vz : uz;
            }
            if (!row_img) continue;
            const float dzp_tzz = ftzz.dplus(g.cz);
            const float dzm_txz = ftxz.dminus(g.cz);
            float term[NV];
            term[0] = -(dt * (bb * bb)) * (ux * (dxm_txx + dzm_txz) + uz * (dzp_tzz + dxp_txz));
            if constexpr (PRECOND) {
                const float dt2 = dt * dt;
                term[1] = dt2 * (wx * wx * (dxm_txx * dxm_txx + dzm_txz * dzm_txz)
                                 + wz * wz * (dzp_tzz * dzp_tzz + dxp_txz * dxp_txz));
            }
            float *q = mine + r * SPB * NV * 64;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if constexpr (ALONE) cell_ref(G + slots[v] * fs, cb) = gc[v] + term[v];
                else q[v * 64] = term[v];
            }
        }
        if constexpr (!ALONE) {
            __syncthreads();
            add_group_terms<NV, SPB>(part[buf], gl, ic, i1, img, lo, hi, G, slots, fs, c0, RB);
        }
    }
}

// ------------------------------------------- adjoint, pyfwi mode, one launch
//
// Both adjoint kernels above in one launch per backward step.  Stage 1 is
// k_adj_str at row i (adjoint stresses, the lambda / mu imaging terms);
// stage 2 is k_adj_vel_grad at row q = i - SDO (adjoint velocities, the rho
// term), fed by stage 1's new stresses through register windows along z and
// lane shifts along x, so the new stresses never go through memory before
// stage 2 reads them.  The arithmetic of every value is that of the two
// kernels.
//
// The adjoint state is read from `ain` and written to `aout` (the caller
// swaps them every step): the tiles around this one read its cells for
// their halo work while it writes.  A wave covers 64 columns and owns the
// 64 - 2*SDO in the middle; its SDO lanes on each side run stage 1 only, for
// the x-derivatives of their neighbours, and each strip runs stage 1 on SDO
// more rows above and below.  Stage 1 is forced to zero off the active cells,
// which the reference never updates.  Per shot and cell the adjoint state is
// read and written once (40 B), the forward state read once (20 B) and the
// model once (12 B): 72 B against 92 B for the two launches, plus the
// accumulators per block's group of shots.  Order 4 only.

template <int SDO, bool PRECOND, bool ALONE, int SPB>
__global__ __launch_bounds__(64 * SPB) void k_adj_fused(
    DevGrid g, const pyfwi_shot *shots, int t, const float *__restrict__ fields,
    int64_t fields_stride, const float *__restrict__ ain, float *__restrict__ aout,
    const float *__restrict__ lam, const float *__restrict__ mu, const float *__restrict__ bu,
    const float *__restrict__ dxp, const float *__restrict__ dzp, float *__restrict__ gacc,
    const float *__restrict__ res)
{
    constexpr int NV1 = PRECOND ? 4 : 2;  // stage 1 terms: mu, lam (, mu_pre, lam_pre)
    constexpr int NV2 = PRECOND ? 2 : 1;  // stage 2 terms: rho (, rho_pre)
    constexpr int NV = NV1 + NV2;
    constexpr int PART = ALONE ? 1 : ADJ_GROUP_ROWS * SPB * NV * 64;
    __shared__ float part[2][PART];
    int s, j, i0, i1;
    bool own, colact;
    if (!tile_fused<SDO, SPB>(g, s, j, i0, i1, own, colact)) return;
    if (!bwd_active(shots, s, t)) return;
    GroupLive gl;  // the shot group (one-shot launches have none)
    if constexpr (!ALONE) gl = group_live<SPB>(g, shots, s, t);
    const int wave = wave_id(), lane = (int)threadIdx.x % 64;
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    // bases 2*SDO rows above cell (0, 0), so that the byte offsets of every
    // row the kernel touches (from -SDO) are non-negative
    const int64_t org = g.origin - (int64_t)2 * SDO * g.pitch;
    const float *AI = ain + (int64_t)s * 5 * fs + org;
    float *AO = aout + (int64_t)s * 5 * fs + org;
    const CellRef avx{AI}, avz{AI + fs}, atxx{AI + 2 * fs}, atzz{AI + 3 * fs}, atxz{AI + 4 * fs};
    const CellRef B{bu + org}, L{lam + org}, M{mu + org};
    const float *F = fields + (int64_t)s * fields_stride + org;
    const CellRef vx{F}, vz{F + fs}, txx{F + 2 * fs}, tzz{F + 3 * fs}, txz{F + 4 * fs};
    float *G = gacc + (int64_t)(s / SPB) * 6 * fs + org;
    constexpr int slots1[4] = {0, 2, 1, 3};  // accumulators of the stage 1 terms
    constexpr int slots2[2] = {4, 5};        // and of the stage 2 terms
    // step t's residual on the adjoint stresses at this lane's receiver
    // (stage 1, halo rows and lanes included); step t-1's on the adjoint
    // velocities it writes (stage 2)
    const RecLine rq = residual_line(shots[s], t, res);
    const int rr = colact && rq.z >= i0 - SDO && rq.z <= i1 + SDO ? rq.index(j) : -1;
    const int rrow = rr >= 0 ? rq.z : -1;
    const float r_txx = rr >= 0 ? rq.base[2 * rq.fs + rr] : 0.0f;
    const float r_tzz = rr >= 0 ? rq.base[3 * rq.fs + rr] : 0.0f;
    const float r_txz = rr >= 0 ? rq.base[4 * rq.fs + rr] : 0.0f;
    const RecLine rq2 = residual_line(shots[s], t - 1, res);
    const int rr2 = own && rq2.z >= i0 && rq2.z <= i1 ? rq2.index(j) : -1;
    const int rrow2 = rr2 >= 0 ? rq2.z : -1;
    const float r_vx = rr2 >= 0 ? rq2.base[rr2] : 0.0f;
    const float r_vz = rr2 >= 0 ? rq2.base[rq2.fs + rr2] : 0.0f;
    const int lo = g.npml + SDO, hi = g.nz - SDO - g.npml, hix = g.nx - SDO - g.npml;
    const bool img = own && j >= lo && j <= hix && i1 >= lo && i0 <= hi;
    const int lo1 = max(i0, lo), hi1 = min(i1, hi);  // imaged rows of the strip
    // a wave with an imaging lane loads the forward state (wave-uniform)
    const bool imgl = __builtin_amdgcn_readfirstlane((int)(__ballot(img) != 0)) != 0;
    const float ex = colact ? dxp[j] : 0.0f;
    const float dt = g.dt;
    auto dx_bvx = [&](uint32_t o) {
        return xplus<SDO>([&](int k) { return avx.at(o, k) * B.at(o, k); }, g.cx); };
    auto dx_bvz = [&](uint32_t o) {
        return xminus<SDO>([&](int k) { return avz.at(o, k) * B.at(o, k); }, g.cx); };

    ZWin<SDO> wbz, wbx;      // b*avz (d/dz-), b*avx (d/dz+) around stage 1's row
    Fifo<SDO> fxp, fxm;      // d/dx+ (b*avx), d/dx- (b*avz), taken on the entering row
    Fifo<2 * SDO> dvx, dvz;  // old adjoint velocities, delayed to stage 2's row
    ZWin<SDO> fvz{};         // forward vz around stage 1's row
    Fifo<SDO> fdvz{};        // its d/dx-, taken on the entering row
    uint32_t cb = ((uint32_t)(i0 - SDO + 2 * SDO) * (uint32_t)g.pitch + (uint32_t)j) * 4u;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        const float bo = B(o), uo = avx(o), wo = avz(o);
        wbz.v[k + 1] = wo * bo;
        wbx.v[k + 1] = uo * bo;
        dvx.v[k] = uo;
        dvz.v[k] = wo;
        if (k >= SDO) {
            fxp.v[k - SDO] = dx_bvx(o);
            fxm.v[k - SDO] = dx_bvz(o);
        }
        if (imgl) {
            Taps<SDO, 1> tvz;
            tvz.load(vz, o);
            fvz.v[k + 1] = tvz(0);
            if (k >= SDO) fdvz.v[k - SDO] = xminus<SDO>(tvz, g.cx);
        }
    }
    ZWin<SDO> wa3{}, wb1{}, wb2{};  // new mu*Atxz (d/dz-), (lam+2mu)*Atzz, lam*Atxx (d/dz+)
    Fifo<SDO> fa1{}, fa2{}, fb3{};  // d/dx- (lam+2mu)*Atxx, d/dx- lam*Atzz, d/dx+ mu*Atxz
    ZWin<SDO> ftzz{}, ftxz{};       // forward tzz (d/dz+), txz (d/dz-) around stage 2's row
    Fifo<SDO> fdtxz{};              // d/dx+ forward txz, delayed to stage 2's row

    int buf = 0;
    for (int ic = i0 - SDO; ic <= i1 + SDO; ic += ADJ_GROUP_ROWS, buf ^= 1) {
        const uint32_t c0 = cb;
        float *mine = &part[buf][(wave * NV) * 64 + lane];
#pragma unroll
        for (int r = 0; r < ADJ_GROUP_ROWS; ++r, cb += RB) {
            const int i = ic + r, q = i - SDO;
            if (i > i1 + SDO) break;
            const uint32_t n = cb + (uint32_t)SDO * RB, cq = cb - (uint32_t)SDO * RB;
            const bool row_img1 = img && i >= lo1 && i <= hi1;
            const bool row_img2 = img && q >= lo1 && q <= hi1;
            // loads: adjoint state, model, forward state, accumulators
            const float bn = B(n), uxn = avx(n), uzn = avz(n);
            const float dxpb_n = dx_bvx(n), dxmb_n = dx_bvz(n);
            const float a0 = atxx(cb), b0 = atzz(cb), d0 = atxz(cb);
            const float l = L(cb), m = M(cb);
            Taps<SDO, 1> tvz;   // forward vz on the entering row (d/dx-)
            Taps<SDO, 0> tvx;   // forward vx on stage 1's row (d/dx+)
            Taps<SDO, 0> ttxz;  // forward txz on stage 1's row (d/dx+)
            Taps<SDO, 1> ttxx;  // forward txx on stage 2's row (d/dx-)
            float f_tzz = 0.0f, s_txx = 0.0f, w_vx = 0.0f, bq = 0.0f;
            if (imgl) {
                tvz.load(vz, n);
                tvx.load(vx, cb);
                ttxz.load(txz, cb);
                ttxx.load(txx, cq);
                f_tzz = tzz(cb);
                if constexpr (PRECOND) {
                    s_txx = txx(cb);
                    w_vx = vx(cq);
                }
                bq = B(cq);
            }
            float gc1[NV1] = {}, gc2[NV2] = {};  // the rows' accumulators (ALONE)
            if constexpr (ALONE) {
                if (row_img2) {
#pragma unroll
                    for (int v = 0; v < NV2; ++v) gc2[v] = cell_ref(G + slots2[v] * fs, cq);
                }
                if (row_img1) {
#pragma unroll
                    for (int v = 0; v < NV1; ++v) gc1[v] = cell_ref(G + slots1[v] * fs, cb);
                }
            }
            // stage 1: adjoint stresses at row i, on every lane
            wbz.push(uzn * bn);
            wbx.push(uxn * bn);
            const float dxpb = fxp.pop(dxpb_n);
            const float dxmb = fxm.pop(dxmb_n);
            const float uxq = dvx.pop(uxn), uzq = dvz.pop(uzn);
            const float ez = dzp[i];
            const float dzmb = wbz.dminus(g.cz);
            const float dzpb = wbx.dplus(g.cz);
            const bool at_rec = i == rrow;
            const float a = at_rec ? a0 + r_txx : a0;
            const float bb = at_rec ? b0 + r_tzz : b0;
            const float d = at_rec ? d0 + r_txz : d0;
            const bool act1 = colact && i >= SDO && i <= g.nz - SDO;
            const float aa = act1 ? fmaf(-dt, ex * a + dxpb, a) : 0.0f;
            const float ab = act1 ? fmaf(-dt, ez * bb + dzmb, bb) : 0.0f;
            const float ad = act1 ? fmaf(-dt, (ex + ez) * d + (dzpb + dxmb), d) : 0.0f;
            // stage 1's products for stage 2: z-windows, and x-derivatives
            // from the neighbouring lanes delayed to stage 2's row
            const float lm = l + 2.0f * m;
            wa3.push(m * ad);
            wb1.push(lm * ab);
            wb2.push(l * aa);
            const float a1 = fa1.pop(xminus<SDO>(LaneNbr<SDO>(lm * aa), g.cx));
            const float a2 = fa2.pop(xminus<SDO>(LaneNbr<SDO>(l * ab), g.cx));
            const float b3 = fb3.pop(xplus<SDO>(LaneNbr<SDO>(m * ad), g.cx));
            // forward state around both rows
            fvz.push(tvz(0));
            const float dxm_vz = fdvz.pop(xminus<SDO>(tvz, g.cx));
            ftzz.push(f_tzz);
            ftxz.push(ttxz(0));
            const float dxp_txz = fdtxz.pop(xplus<SDO>(ttxz, g.cx));
            float term1[NV1], term2[NV2];
            if (row_img1) {
                const float dxp_vx = xplus<SDO>(tvx, g.cx);
                const float dzp_vz = fvz.dplus(g.cz);
                const float dzm_vz = fvz.dminus(g.cz);
                term1[0] = 2.0f * dt * (aa * dxp_vx + ab * dzm_vz) + dt * ad * (dzp_vz + dxm_vz);
                term1[1] = dt * (aa + ab) * (dxp_vx + dzm_vz);
                if constexpr (PRECOND) {
                    const float dt2 = dt * dt;
                    const float s_tzz = ftzz.v[2 * SDO], s_txz = ftxz.v[2 * SDO];  // row i
                    term1[2] = 4.0f * dt2 * (s_txx * s_txx * dxp_vx * dxp_vx +
                                             s_tzz * s_tzz * dzm_vz * dzm_vz) +
                               dt2 * s_txz * s_txz * (dzp_vz * dzp_vz + dxm_vz * dxm_vz);
                    term1[3] = dt2 * (s_txx * s_txx + s_tzz * s_tzz) *
                               (dxp_vx * dxp_vx + dzm_vz * dzm_vz);
                }
            }
            // stage 2: adjoint velocities at row q
            const float a3 = wa3.dminus(g.cz);
            const float b1 = wb1.dplus(g.cz);
            const float b2 = wb2.dplus(g.cz);
            const float ezq = dzp[q];
            const float ux = fmaf(-dt, (2.0f * ex + ezq) * uxq + (a1 + a2 + a3), uxq);
            const float uz = fmaf(-dt, (ex + 2.0f * ezq) * uzq + (b2 + b1 + b3), uzq);
            if (row_img2) {
                const float dxm_txx = xminus<SDO>(ttxx, g.cx);
                const float dzp_tzz = ftzz.dplus(g.cz);
                const float dzm_txz = ftxz.dminus(g.cz);
                term2[0] = -(dt * (bq * bq)) * (ux * (dxm_txx + dzm_txz) + uz * (dzp_tzz + dxp_txz));
                if constexpr (PRECOND) {
                    const float dt2 = dt * dt;
                    const float wz = fvz.v[0];  // forward vz at row q
                    term2[1] = dt2 * (w_vx * w_vx * (dxm_txx * dxm_txx + dzm_txz * dzm_txz) +
                                      wz * wz * (dzp_tzz * dzp_tzz + dxp_txz * dxp_txz));
                }
            }
            // stores
            if (own && i >= i0 && i <= i1) {
                cell_ref(AO + 2 * fs, cb) = aa;
                cell_ref(AO + 3 * fs, cb) = ab;
                cell_ref(AO + 4 * fs, cb) = ad;
            }
            if (own && q >= i0) {
                const bool at_rec2 = q == rrow2;
                cell_ref(AO, cq) = at_rec2 ? ux + r_vx : ux;
                cell_ref(AO + fs, cq) = at_rec2 ? uz + r_vz : uz;
            }
            if constexpr (ALONE) {
                if (row_img1) {
#pragma unroll
                    for (int v = 0; v < NV1; ++v) cell_ref(G + slots1[v] * fs, cb) = gc1[v] + term1[v];
                }
                if (row_img2) {
#pragma unroll
                    for (int v = 0; v < NV2; ++v) cell_ref(G + slots2[v] * fs, cq) = gc2[v] + term2[v];
                }
            } else {
                float *pq = mine + r * SPB * NV * 64;
                if (row_img1) {
#pragma unroll
                    for (int v = 0; v < NV1; ++v) pq[v * 64] = term1[v];
                }
                if (row_img2) {
#pragma unroll
                    for (int v = 0; v < NV2; ++v) pq[(NV1 + v) * 64] = term2[v];
                }
            }
        }
        if constexpr (!ALONE) {
            __syncthreads();
            // the staged rows dealt out to the live waves, each adding the
            // shots' terms in shot order onto the group's accumulators
            for (int r = gl.rank; r < ADJ_GROUP_ROWS; r += gl.count) {
                const int i = ic + r, q = i - SDO;
                if (i > i1 + SDO) break;
                if (!img) continue;
                const uint32_t o1 = c0 + (uint32_t)r * RB, o2 = o1 - (uint32_t)SDO * RB;
                const float *pr = &part[buf][r * SPB * NV * 64 + lane];
                if (i >= lo1 && i <= hi1) {
                    float cur[NV1], acc[NV1];
#pragma unroll
                    for (int v = 0; v < NV1; ++v) {
                        cur[v] = cell_ref(G + slots1[v] * fs, o1);
                        acc[v] = 0.0f;
                    }
#pragma unroll
                    for (int w = 0; w < SPB; ++w) {
                        if (!(gl.mask >> w & 1u)) continue;
#pragma unroll
                        for (int v = 0; v < NV1; ++v) acc[v] += pr[(w * NV + v) * 64];
                    }
#pragma unroll
                    for (int v = 0; v < NV1; ++v) cell_ref(G + slots1[v] * fs, o1) = cur[v] + acc[v];
                }
                if (q >= lo1 && q <= hi1) {
                    float cur[NV2], acc[NV2];
#pragma unroll
                    for (int v = 0; v < NV2; ++v) {
                        cur[v] = cell_ref(G + slots2[v] * fs, o2);
                        acc[v] = 0.0f;
                    }
#pragma unroll
                    for (int w = 0; w < SPB; ++w) {
                        if (!(gl.mask >> w & 1u)) continue;
#pragma unroll
                        for (int v = 0; v < NV2; ++v) acc[v] += pr[(w * NV + NV1 + v) * 64];
                    }
#pragma unroll
                    for (int v = 0; v < NV2; ++v) cell_ref(G + slots2[v] * fs, o2) = cur[v] + acc[v];
                }
            }
        }
    }
}

// ------------------------------------------------- adjoint, exact mode
//
// Exact transpose of the forward step, with D+^T = -D- and D-^T = -D+ on
// fields that vanish outside the active cells, and the damping factor
// keep = 1 - dt*(dx+dz) transposing to itself.  Per backward step t:
//   k_exact_str: q = Atau*keep (taps);  Av += -dt*D(C q);  g_lam, g_mu
//   k_exact_vel: p = Av*keep (taps);  Atau = q - dt*D(b p);  Av' = p;  g_b;
//                then the receiver sampling of S_t adds the residual row.
// The velocity adjoint ping-pongs between two buffers because its taps are
// read while its centre is written.

// vel: [nshots][2][fs] velocities after the velocity half step of step t
// (V_t); gacc: [adj_acc_sets(sdo, nshots)][6][fs] with slots 0 (mu) and 2
// (lam) updated, the shots of a block summing their terms first as in the
// reference mode (every active cell here).  Column marching as the other
// stencil kernels: z-windows of the products differentiated along z,
// x-derivatives on the entering row delayed SDO rows.
template <int SDO, bool ALONE, int SPB>
__global__ __launch_bounds__(64 * SPB) void k_exact_str(
    DevGrid g, const pyfwi_shot *shots, int t, const float *vel, int64_t vel_shot_stride,
    float *adj, float *av, const float *__restrict__ lam,
    const float *__restrict__ mu, const float *__restrict__ dxp,
    const float *__restrict__ dzp, float *gacc)
{
    constexpr int NV = 2;  // staged terms: mu, lam
    __shared__ float part[2][ADJ_GROUP_ROWS * SPB * NV * 64];
    int s, j, i0, i1;
    if (!tile<SDO, SPB>(g, s, j, i0, i1)) return;
    if (t >= shots[s].nt) return;
    GroupLive gl;  // the shot group (one-shot launches have none)
    if constexpr (!ALONE) gl = group_live<SPB, true>(g, shots, s, t);
    const int wave = wave_id(), lane = (int)threadIdx.x % 64;
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    const float dt = g.dt;
    float *A = adj + (int64_t)s * 5 * fs + g.origin;
    float *Av = av + (int64_t)s * 2 * fs + g.origin;
    const CellRef qxx{A + 2 * fs}, qzz{A + 3 * fs}, qxz{A + 4 * fs};
    const CellRef L{lam + g.origin}, M{mu + g.origin};
    const float *V = vel + (int64_t)s * vel_shot_stride + g.origin;
    const CellRef vx{V}, vz{V + fs};
    float *G = gacc + (int64_t)(s / SPB) * 6 * fs + g.origin;
    constexpr int slots[2] = {0, 2};  // accumulator of each staged term
    // keep = 1 - dt*(dx + dz) of row r at column offset k
    auto kp = [&](int r, int k) { return 1.0f - dt * (dxp[j + k] + dzp[r]); };
    ZWin<SDO> wa3, wb1, wb2;  // mu*qxz*keep (d/dz-), lam*qxx*keep, (lam+2mu)*qzz*keep (d/dz+)
    ZWin<SDO> wvz, wvx;       // forward vz (d/dz-), vx (d/dz+)
    Fifo<SDO> fa1, fa2, fa3, fvx, fvz;  // x-derivatives of the entering row
    auto x_a1 = [&](uint32_t o, int r) {
        return xminus<SDO>([&](int k) {
            return (L.at(o, k) + 2.0f * M.at(o, k)) * (qxx.at(o, k) * kp(r, k)); }, g.cx); };
    auto x_a2 = [&](uint32_t o, int r) {
        return xminus<SDO>([&](int k) { return L.at(o, k) * (qzz.at(o, k) * kp(r, k)); }, g.cx); };
    auto x_a3 = [&](uint32_t o, int r) {
        return xplus<SDO>([&](int k) { return M.at(o, k) * (qxz.at(o, k) * kp(r, k)); }, g.cx); };
    auto x_vx = [&](uint32_t o) { return xplus<SDO>([&](int k) { return vx.at(o, k); }, g.cx); };
    auto x_vz = [&](uint32_t o) { return xminus<SDO>([&](int k) { return vz.at(o, k); }, g.cx); };
    auto z_a3 = [&](uint32_t o, int r) { return M(o) * (qxz(o) * kp(r, 0)); };
    auto z_b1 = [&](uint32_t o, int r) { return L(o) * (qxx(o) * kp(r, 0)); };
    auto z_b2 = [&](uint32_t o, int r) { return (L(o) + 2.0f * M(o)) * (qzz(o) * kp(r, 0)); };
    uint32_t cb = ((uint32_t)i0 * (uint32_t)g.pitch + (uint32_t)j) * 4u;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        const int r = i0 + k - SDO;
        wa3.v[k + 1] = z_a3(o, r);
        wb1.v[k + 1] = z_b1(o, r);
        wb2.v[k + 1] = z_b2(o, r);
        wvz.v[k + 1] = vz(o);
        wvx.v[k + 1] = vx(o);
        if (k >= SDO) {
            const int q = k - SDO;
            fa1.v[q] = x_a1(o, r);
            fa2.v[q] = x_a2(o, r);
            fa3.v[q] = x_a3(o, r);
            fvx.v[q] = x_vx(o);
            fvz.v[q] = x_vz(o);
        }
    }
    // ALONE (launches of one shot): the shot adds its terms itself
    int buf = 0;
    for (int ic = i0; ic <= i1; ic += ADJ_GROUP_ROWS, buf ^= 1) {
        const uint32_t c0 = cb;
        float *mine = &part[buf][wave * NV * 64 + lane];
#pragma unroll MARCH_UNROLL
        for (int r = 0; r < ADJ_GROUP_ROWS; ++r, cb += RB) {
            const int i = ic + r;
            if (i > i1) break;
            const uint32_t n = cb + (uint32_t)SDO * RB;
            const int rn = i + SDO;
            wa3.push(z_a3(n, rn));
            wb1.push(z_b1(n, rn));
            wb2.push(z_b2(n, rn));
            wvz.push(vz(n));
            wvx.push(vx(n));
            const float a1 = fa1.pop(x_a1(n, rn)), a2 = fa2.pop(x_a2(n, rn));
            const float a3x = fa3.pop(x_a3(n, rn));
            const float exx = fvx.pop(x_vx(n)), gzx = fvz.pop(x_vz(n));
            float *pax = &cell_ref(Av, cb), *paz = &cell_ref(Av + fs, cb);
            const float ax0 = *pax, az0 = *paz;
            const float keep = kp(i, 0);
            const float q1 = qxx(cb) * keep, q2 = qzz(cb) * keep, q3 = qxz(cb) * keep;
            float gc[NV] = {};  // the row's accumulators (ALONE)
            if constexpr (ALONE) {
#pragma unroll
                for (int v = 0; v < NV; ++v) gc[v] = cell_ref(G + slots[v] * fs, cb);
            }
            const float sx = a1 + a2 + wa3.dminus(g.cz);
            const float sz = wb1.dplus(g.cz) + wb2.dplus(g.cz) + a3x;
            *pax = fmaf(-dt, sx, ax0);
            *paz = fmaf(-dt, sz, az0);
            const float ezz = wvz.dminus(g.cz);
            const float gxz = gzx + wvx.dplus(g.cz);
            float term[NV];
            term[0] = dt * (2.0f * q1 * exx + 2.0f * q2 * ezz + q3 * gxz);
            term[1] = dt * (q1 + q2) * (exx + ezz);
            float *q = mine + r * SPB * NV * 64;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                if constexpr (ALONE) cell_ref(G + slots[v] * fs, cb) = gc[v] + term[v];
                else q[v * 64] = term[v];
            }
        }
        if constexpr (!ALONE) {
            __syncthreads();
            add_group_terms<NV, SPB>(part[buf], gl, ic, i1, true, i0, i1, G, slots, fs, c0, RB);
        }
    }
}

// tau: [nshots][3][fs] stresses after the injection of step t (input of the
// update); gacc slot 4 accumulates dJ/db (per block's group of shots, as in
// k_exact_str).  Writes the new stress adjoint in place and p = Av*keep to
// av_out.  (The receiver sampling of S_t, which adds the residual row, is
// k_exact_inject's.)
template <int SDO, bool ALONE, int SPB>
__global__ __launch_bounds__(64 * SPB) void k_exact_vel(
    DevGrid g, const pyfwi_shot *shots, int t, const float *tau, int64_t tau_shot_stride,
    float *adj, const float *av_in, float *av_out, const float *__restrict__ bu,
    const float *__restrict__ dxp, const float *__restrict__ dzp, float *gacc,
    const float *__restrict__ src)
{
    constexpr int NV = 1;  // staged term: b
    __shared__ float part[2][ADJ_GROUP_ROWS * SPB * NV * 64];
    int s, j, i0, i1;
    if (!tile<SDO, SPB>(g, s, j, i0, i1)) return;
    if (t >= shots[s].nt) return;
    GroupLive gl;  // the shot group (one-shot launches have none)
    if constexpr (!ALONE) gl = group_live<SPB, true>(g, shots, s, t);
    const int wave = wave_id(), lane = (int)threadIdx.x % 64;
    const int64_t fs = g.fstride;
    const uint32_t RB = (uint32_t)g.pitch * 4u;
    const float dt = g.dt;
    float *A = adj + (int64_t)s * 5 * fs + g.origin;
    const float *Ain = av_in + (int64_t)s * 2 * fs + g.origin;
    float *Aout = av_out + (int64_t)s * 2 * fs + g.origin;
    const CellRef pvx{Ain}, pvz{Ain + fs}, B{bu + g.origin};
    const float *T = tau + (int64_t)s * tau_shot_stride + g.origin;
    const CellRef txx{T}, tzz{T + fs}, txz{T + 2 * fs};
    float *G = gacc + (int64_t)(s / SPB) * 6 * fs + g.origin;
    constexpr int slots[1] = {4};  // accumulator of the staged term
    auto kp = [&](int r, int k) { return 1.0f - dt * (dxp[j + k] + dzp[r]); };
    ZWin<SDO> wbz, wbx;    // b*pvz*keep (d/dz-), b*pvx*keep (d/dz+)
    ZWin<SDO> wtxz, wtzz;  // forward txz (d/dz-), tzz (d/dz+)
    Fifo<SDO> frxx, frxz, ftxx, ftxz;  // x-derivatives of the entering row
    auto x_bx = [&](uint32_t o, int r) {
        return xplus<SDO>([&](int k) { return B.at(o, k) * (pvx.at(o, k) * kp(r, k)); }, g.cx); };
    auto x_bz = [&](uint32_t o, int r) {
        return xminus<SDO>([&](int k) { return B.at(o, k) * (pvz.at(o, k) * kp(r, k)); }, g.cx); };
    // With `src` the stresses come from a kept state, which holds no source
    // term: step t's is added to txx and tzz as they are loaded at the
    // source cell, giving the values a recomputed segment holds in memory
    // (k_inject_record adds them there before the velocity update).
    SrcTerm st = src ? src_term(shots[s], t, src) : SrcTerm{};
    if (st.z < i0 - SDO || st.z > i1 + SDO || st.x < j - SDO || st.x > j + SDO) st.z = -1;
    auto x_txx = [&](uint32_t o, int r) {
        return xminus<SDO>([&](int k) {
            const float v = txx.at(o, k);
            return r == st.z && j + k == st.x ? v + st.v[2] : v; }, g.cx); };
    auto tzz_at = [&](uint32_t o, int r) {
        const float v = tzz(o);
        return r == st.z && j == st.x ? v + st.v[3] : v; };
    auto x_txz = [&](uint32_t o) { return xplus<SDO>([&](int k) { return txz.at(o, k); }, g.cx); };
    auto z_bz = [&](uint32_t o, int r) { return B(o) * (pvz(o) * kp(r, 0)); };
    auto z_bx = [&](uint32_t o, int r) { return B(o) * (pvx(o) * kp(r, 0)); };
    uint32_t cb = ((uint32_t)i0 * (uint32_t)g.pitch + (uint32_t)j) * 4u;
#pragma unroll
    for (int k = 0; k < 2 * SDO; ++k) {
        const uint32_t o = cb + (uint32_t)(k - SDO) * RB;
        const int r = i0 + k - SDO;
        wbz.v[k + 1] = z_bz(o, r);
        wbx.v[k + 1] = z_bx(o, r);
        wtxz.v[k + 1] = txz(o);
        wtzz.v[k + 1] = tzz_at(o, r);
        if (k >= SDO) {
            const int q = k - SDO;
            frxx.v[q] = x_bx(o, r);
            frxz.v[q] = x_bz(o, r);
            ftxx.v[q] = x_txx(o, r);
            ftxz.v[q] = x_txz(o);
        }
    }
    int buf = 0;
    for (int ic = i0; ic <= i1; ic += ADJ_GROUP_ROWS, buf ^= 1) {
        const uint32_t c0 = cb;
        float *mine = &part[buf][wave * NV * 64 + lane];
#pragma unroll MARCH_UNROLL
        for (int r = 0; r < ADJ_GROUP_ROWS; ++r, cb += RB) {
            const int i = ic + r;
            if (i > i1) break;
            const uint32_t n = cb + (uint32_t)SDO * RB;
            const int rn = i + SDO;
            wbz.push(z_bz(n, rn));
            wbx.push(z_bx(n, rn));
            wtxz.push(txz(n));
            wtzz.push(tzz_at(n, rn));
            const float rxx = frxx.pop(x_bx(n, rn)), rxz_x = frxz.pop(x_bz(n, rn));
            const float dtx = ftxx.pop(x_txx(n, rn)), dtz = ftxz.pop(x_txz(n));
            const float keep = kp(i, 0);
            const float px = pvx(cb) * keep, pz = pvz(cb) * keep;
            float *pxx = &cell_ref(A + 2 * fs, cb), *pzz = &cell_ref(A + 3 * fs, cb),
                  *pxz = &cell_ref(A + 4 * fs, cb);
            const float axx = *pxx, azz = *pzz, axz = *pxz;
            float gc = 0.0f;  // the row's accumulator (ALONE)
            if constexpr (ALONE) gc = cell_ref(G + 4 * fs, cb);
            const float rzz = wbz.dminus(g.cz);
            const float rxz = wbx.dplus(g.cz) + rxz_x;
            *pxx = fmaf(-dt, rxx, axx * keep);
            *pzz = fmaf(-dt, rzz, azz * keep);
            *pxz = fmaf(-dt, rxz, axz * keep);
            cell_ref(Aout, cb) = px;
            cell_ref(Aout + fs, cb) = pz;
            const float dtau_x = dtx + wtxz.dminus(g.cz);
            const float dtau_z = wtzz.dplus(g.cz) + dtz;
            const float term = dt * (px * dtau_x + pz * dtau_z);
            if constexpr (ALONE) cell_ref(G + 4 * fs, cb) = gc + term;
            else mine[r * SPB * NV * 64] = term;
        }
        if constexpr (!ALONE) {
            __syncthreads();
            add_group_terms<NV, SPB>(part[buf], gl, ic, i1, true, i0, i1, G, slots, fs, c0, RB);
        }
    }
}

// Residual of sample t/dt_scale added at the receivers after backward step t
// (adjoint of the gather of S_t: component f goes to the field it sampled).
// Velocity components go to the ping-pong buffer `av`, stresses to
// adj[2..4].
__global__ void k_exact_inject(DevGrid g, const pyfwi_shot *shots, int t, float *adj,
                               float *av, const float *__restrict__ res,
                               const int32_t *__restrict__ rtab)
{
    const int s = blockIdx.x;
    const pyfwi_shot sh = shots[s];
    if (t >= sh.nt || (t % sh.dt_scale) != 0) return;
    const int row = t / sh.dt_scale;
    if (row >= sh.n_samples) return;
    const int64_t fs = g.fstride;
    float *A = adj + (int64_t)s * 5 * fs + g.origin;
    float *V = av + (int64_t)s * 2 * fs + g.origin;
    const float *rr = res + sh.rec_offset;
    auto target = [&](int f, int64_t c) {
        const int d = recorded_field(sh, f);
        return d < 2 ? &V[d * fs + c] : &A[d * fs + c];
    };
    if (sh.rec_layout == PYFWI_REC_TABLE) {
        const int32_t *cells = rtab + sh.rec_table;
        for (int r = threadIdx.x; r < sh.n_rec; r += blockDim.x) {
            const int32_t e = cells[r];
            if (e < 0) continue;
            const int64_t c = table_cell(g, e);
            for (int f = 0; f < 5; ++f)
                atomicAdd(target(f, c), rr[((int64_t)f * sh.n_samples + row) * sh.rec_stride + r]);
        }
        return;
    }
    const int64_t c0 = (int64_t)sh.rec_z * g.pitch + sh.rec_x0;
    for (int r = threadIdx.x; r < sh.n_rec; r += blockDim.x) {
        const int64_t c = c0 + (int64_t)r * sh.rec_dx;
        for (int f = 0; f < 5; ++f)
            *target(f, c) += rr[((int64_t)f * sh.n_samples + row) * sh.rec_stride + r];
    }
}

// ------------------------------------------------------------ L2 misfit
//
// Residual of the records of the shots of a pass against observed records
// (CostFunction.l2, src/PyFWI/fwi_tools.py:1402-1422): res = seis - obs in
// FP32 for the selected components, misfit = 0.5 * sum(res^2) accumulated in
// FP64.  comp_mask bit f selects component f (vx, vz, txx, tzz, txz); mask 0
// means the pressure output (txx + tzz)/2 compared with obs[txx], whose
// residual goes to both normal stresses and counts twice, as the reference's
// 'taux'/'tauz' pair does.  obs == nullptr: observed records of zero;
// misfit == nullptr: the residual only.
__global__ void k_l2_residual(const pyfwi_shot *shots, const float *__restrict__ seis,
                              const float *__restrict__ obs, float *res, int comp_mask,
                              double *misfit)
{
    const int s = blockIdx.y;
    const pyfwi_shot sh = shots[s];
    const int64_t plane = (int64_t)sh.n_samples * sh.rec_stride;
    const float *S = seis + sh.rec_offset, *D = obs ? obs + sh.rec_offset : nullptr;
    float *R = res + sh.rec_offset;
    double acc = 0.0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < plane;
         k += (int64_t)gridDim.x * blockDim.x) {
        if ((int)(k % sh.rec_stride) >= sh.n_rec) {
#pragma unroll
            for (int f = 0; f < 5; ++f) R[f * plane + k] = 0.0f;
            continue;
        }
        if (comp_mask == 0) {
            const float p = (S[2 * plane + k] + S[3 * plane + k]) / 2;
            const float r = D ? p - D[2 * plane + k] : p;
            R[k] = R[plane + k] = R[4 * plane + k] = 0.0f;
            R[2 * plane + k] = r;
            R[3 * plane + k] = r;
            acc += (double)r * r;
        } else {
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                float r = 0.0f;
                if (comp_mask & (1 << f)) {
                    r = D ? S[f * plane + k] - D[f * plane + k] : S[f * plane + k];
                    acc += 0.5 * (double)r * r;
                }
                R[f * plane + k] = r;
            }
        }
    }
    if (!misfit) return;
    // block reduction, one atomic per block
    __shared__ double part[256];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(misfit, part[0]);
}

void launch_l2_residual(int nshots, const pyfwi_shot *shots, const float *seis, const float *obs,
                        float *res, int comp_mask, double *misfit, hipStream_t st)
{
    hipLaunchKernelGGL(k_l2_residual, dim3(64, nshots), dim3(256), 0, st, shots, seis, obs, res,
                       comp_mask, misfit);
}

// ------------------------------------------------------ record filtering
//
// The multiscale lowpass of the misfit (lowpass / adj_lowpass,
// src/PyFWI/fwi_tools.py:1189-1255) pads each trace of n samples with `pad`
// zeros, multiplies its N-point spectrum (N = n + pad) by H and conj(H) (or
// conj(H) and H) and keeps the first n samples.  |H|^2 is real and even on
// the N-point circle, so both are the same circular convolution with the
// symmetric kernel k = ifft(|H|^2):
//   y[t] = sum_{m=-L..L} k[|m|] x[(t - m) mod N],  x = 0 past the trace,
// with the taps k[0..L] computed on the host (L: where k has decayed).  A
// block filters 64 time samples of 64 traces (one receiver each lane) of one
// shot and component: 4 waves, 16 consecutive outputs per lane; the input
// window [t0 - L, t0 + 64 + L) is staged through LDS 64 rows at a time, and
// the taps of a lane's 16 outputs slide through registers one tap per row.

constexpr int FIR_T = 64;      // time samples per block
constexpr int FIR_C = 64;      // window rows per LDS chunk
constexpr int FIR_J = 16;      // outputs per lane (FIR_T / 4 waves)

__global__ __launch_bounds__(256) void k_fir(const pyfwi_shot *shots, const float *__restrict__ in,
                                             float *__restrict__ out, int comp_mask,
                                             const float *__restrict__ taps, int L, int pad)
{
    extern __shared__ float fir_lds[];
    const int s = blockIdx.y / 5, f = blockIdx.y % 5;
    const pyfwi_shot sh = shots[s];
    const bool on = comp_mask == 0 ? (f == 2 || f == 3) : ((comp_mask >> f) & 1);
    const int ns = sh.n_samples, nr = sh.n_rec;
    const int t0 = blockIdx.x * FIR_T, r0 = blockIdx.z * 64;
    if (!on || t0 >= ns || r0 >= nr) return;  // whole block: uniform
    const int N = ns + pad;
    const int lane = threadIdx.x % 64, tg = threadIdx.x / 64;
    const int r = r0 + lane;
    const int64_t plane = (int64_t)ns * sh.rec_stride;
    const float *X = in + sh.rec_offset + f * plane;
    float *Y = out + sh.rec_offset + f * plane;
    // taps k[m], m = -(L + FIR_T + FIR_C) .. L + FIR_T + FIR_C, zero past +-L
    const int M = L + FIR_T + FIR_C;
    float *ks = fir_lds;                    // 2M + 1 values, ks[M + m] = k[|m|]
    float *xs = fir_lds + ((2 * M + 1 + 63) / 64) * 64;  // FIR_C x 64 window rows
    for (int k = threadIdx.x; k < 2 * M + 1; k += 256) {
        const int m = k - M;
        const int a = m < 0 ? -m : m;
        ks[k] = a <= L ? taps[a] : 0.0f;
    }
    float acc[FIR_J];
#pragma unroll
    for (int j = 0; j < FIR_J; ++j) acc[j] = 0.0f;
    const int tj0 = t0 + tg * FIR_J;  // this lane's first output
    const int w0 = t0 - L, w1 = t0 + FIR_T + L;  // input window [w0, w1)
    for (int c0 = w0; c0 < w1; c0 += FIR_C) {
        __syncthreads();  // taps staged / previous chunk consumed
        for (int k = threadIdx.x; k < FIR_C * 64; k += 256) {
            const int row = k / 64, col = k % 64;
            int idx = c0 + row;  // circular index on N, zero past the trace
            if (idx < 0) idx += N;
            if (idx >= N) idx -= N;
            const int rr = r0 + col;
            float v = 0.0f;
            if (c0 + row < w1 && idx >= 0 && idx < ns && rr < nr)
                v = X[(int64_t)idx * sh.rec_stride + rr];
            xs[row * 64 + col] = v;
        }
        __syncthreads();
        // kw[j] = k[tj0 + j - tau] for tau = c0 + c (sliding one tap per row)
        float kw[FIR_J];
        const int base = M + tj0 - c0;
#pragma unroll
        for (int j = 0; j < FIR_J; ++j) kw[j] = ks[base + j];
        for (int c = 0; c < FIR_C; ++c) {
            const float xv = xs[c * 64 + lane];
#pragma unroll
            for (int j = 0; j < FIR_J; ++j) acc[j] = fmaf(kw[j], xv, acc[j]);
#pragma unroll
            for (int j = FIR_J - 1; j > 0; --j) kw[j] = kw[j - 1];
            kw[0] = ks[base - c - 1];
        }
    }
    if (r >= nr) return;
#pragma unroll
    for (int j = 0; j < FIR_J; ++j)
        if (tj0 + j < ns) Y[(int64_t)(tj0 + j) * sh.rec_stride + r] = acc[j];
}

void launch_fir(int nshots, const pyfwi_shot *shots, const pyfwi_shot *h_shots, const float *in,
                float *out, int comp_mask, const float *taps, int L, int pad, hipStream_t st)
{
    int ns_max = 0, nr_max = 0;
    for (int k = 0; k < nshots; ++k) {
        ns_max = std::max(ns_max, h_shots[k].n_samples);
        nr_max = std::max(nr_max, h_shots[k].n_rec);
    }
    if (ns_max == 0 || nr_max == 0) return;
    const int M = L + FIR_T + FIR_C;
    const size_t lds = (((size_t)(2 * M + 1) + 63) / 64 * 64 + (size_t)FIR_C * 64) * sizeof(float);
    dim3 grid((ns_max + FIR_T - 1) / FIR_T, nshots * 5, (nr_max + 63) / 64);
    hipLaunchKernelGGL(k_fir, grid, dim3(256), lds, st, shots, in, out, comp_mask, taps, L, pad);
}

// ------------------------------------------------------------- utilities

// Copy `count` floats per shot (count % 4 == 0, 16-byte aligned bases).
// Zero n floats from p (p 16-byte aligned): a grid-stride fill of 16-byte
// stores, which clears the large per-pass buffers at close to copy speed.
__global__ void k_zero(float *p, int64_t n)
{
    const int64_t n4 = n / 4;
    float4 *q = reinterpret_cast<float4 *>(p);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n4; k += stride)
        q[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int64_t k = 4 * n4 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += stride)
        p[k] = 0.0f;
}

// Zero for buffers off the 16-byte grid (no memset node, so the clear can
// sit in a captured graph).
__global__ void k_zero1(float *p, int64_t n)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        p[k] = 0.0f;
}

__global__ void k_copy_shots(float *dst, int64_t dst_stride, const float *src,
                             int64_t src_stride, int64_t count)
{
    const int s = blockIdx.y;
    const float4 *a = reinterpret_cast<const float4 *>(src + (int64_t)s * src_stride);
    float4 *b = reinterpret_cast<float4 *>(dst + (int64_t)s * dst_stride);
    const int64_t n4 = count / 4;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n4;
         k += (int64_t)gridDim.x * blockDim.x)
        b[k] = a[k];
}

// Sum `nsets` accumulator sets (one per shot in exact mode, one per shot
// group in reference mode) into the unpadded output [nslot][tnz*tnx]; exact
// mode converts dJ/db (slot 4) to dJ/drho = -b^2 dJ/db.  With `accumulate`
// the sets are added one by one onto the running output (the converted slot
// adds its pass total instead).
__global__ void k_grad_sum(DevGrid g, int nsets, int nslot, const float *gacc,
                           float *out, const float *__restrict__ bu, int b_to_rho, int accumulate)
{
    const int64_t n = (int64_t)g.nz * g.nx;
    const int64_t fs = g.fstride;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n * nslot;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int slot = (int)(k / n);
        const int64_t cell = k - (int64_t)slot * n;
        const int i = (int)(cell / g.nx), j = (int)(cell - (int64_t)i * g.nx);
        const int64_t p = g.origin + (int64_t)i * g.pitch + j;
        const bool conv = b_to_rho && slot == 4;
        float acc = accumulate && !conv ? out[k] : 0.0f;
        for (int s = 0; s < nsets; ++s) acc += gacc[((int64_t)s * 6 + slot) * fs + p];
        if (conv) {
            acc = -(bu[p] * bu[p]) * acc;
            if (accumulate) acc += out[k];
        }
        out[k] = acc;
    }
}

// Scatter a tnz x tnx array into the padded layout (padding untouched).
__global__ void k_pad(DevGrid g, const float *src, float *dst)
{
    const int64_t n = (int64_t)g.nz * g.nx;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(k / g.nx), j = (int)(k - (int64_t)i * g.nx);
        dst[g.origin + (int64_t)i * g.pitch + j] = src[k];
    }
}

// Model check over the padded model: out[0] = bits of max((lam+2mu)*b), the
// squared largest P velocity (non-negative floats order like their bit
// patterns); out[1] = 1 when lam, mu or b holds NaN or Inf.  out: two ints
// zeroed by the caller.
__global__ void k_model_check(DevGrid g, const float *__restrict__ lam,
                              const float *__restrict__ mu, const float *__restrict__ bu,
                              int *out)
{
    const int64_t n = (int64_t)g.nz * g.nx;
    float m = 0.0f;
    int bad = 0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(k / g.nx), j = (int)(k - (int64_t)i * g.nx);
        const int64_t c = g.origin + (int64_t)i * g.pitch + j;
        const float l = lam[c], u = mu[c], b = bu[c];
        if (!isfinite(l) || !isfinite(u) || !isfinite(b)) bad = 1;
        else m = fmaxf(m, (l + 2.0f * u) * b);
    }
    // wave-wide reduction, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        m = fmaxf(m, __shfl_xor(m, off, 64));
        bad |= __shfl_xor(bad, off, 64);
    }
    if (threadIdx.x % 64 == 0) {
        atomicMax(&out[0], __float_as_int(fmaxf(m, 0.0f)));
        if (bad) atomicOr(&out[1], 1);
    }
}

// Gather the padded layout back to tnz x tnx, `nf` consecutive fields.
__global__ void k_unpad(DevGrid g, const float *src, float *dst, int nf)
{
    const int64_t n = (int64_t)g.nz * g.nx;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n * nf;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int f = (int)(k / n);
        const int64_t cell = k - (int64_t)f * n;
        const int i = (int)(cell / g.nx), j = (int)(cell - (int64_t)i * g.nx);
        dst[k] = src[(int64_t)f * g.fstride + g.origin + (int64_t)i * g.pitch + j];
    }
}

// ------------------------------------------------------- launch wrappers


// Strip height for a launch of `per_strip` working wavefronts per row strip:
// as many strips as keep the launch within `per_cu` such waves per CU (tall
// strips, fewer halo rows, when many shots are batched; short strips when a
// single shot must still fill the GPU), within [min_rows, max_rows].  Waves
// of a block without a shot leave at once and are not counted.  The height is
// then evened out over the strips, so the last strip is not a sliver that
// costs a round of waves of its own.
static int strip_rows(const DevGrid &g, int nz_act, int64_t per_strip, int per_cu, int min_rows,
                      int max_rows)
{
    if (g.rows_override > 0) return g.rows_override;
    const int64_t room = (int64_t)per_cu * g.ncu;
    const int nstrips = (int)std::max<int64_t>(1, std::min<int64_t>(nz_act, room / per_strip));
    int rows = (nz_act + nstrips - 1) / nstrips;
    rows = std::max(min_rows, std::min(max_rows, rows));
    const int n = (nz_act + rows - 1) / rows;
    return std::max(min_rows, (nz_act + n - 1) / n);
}

// Strip height of a column-marching launch of `nshots` shots, `ncb` column
// blocks each (tall_ok: the exact adjoint's kernels, whose launches of 20
// shots at 540 x 1040 cells ran 4-7 % faster with the one-round 36-row strips
// than with 12, 18 or 24 rows).
static int march_rows(const DevGrid &g, int nz_act, int ncb, int nshots, bool tall_ok = false)
{
    if (nshots == 1)
        return strip_rows(g, nz_act, ncb, MARCH_WAVES_PER_CU_ONE_SHOT, MARCH_ROWS_MIN_ONE_SHOT,
                          MARCH_ROWS_MAX_ONE_SHOT);
    const int one_round = strip_rows(g, nz_act, (int64_t)ncb * nshots, MARCH_WAVES_PER_CU,
                                     MARCH_ROWS_MIN, MARCH_ROWS_MAX);
    if (tall_ok || one_round <= MARCH_ROWS_ONE_ROUND || g.rows_override > 0) return one_round;
    return strip_rows(g, nz_act, 1, 1, MARCH_ROWS_MANY, MARCH_ROWS_MANY);
}

// Shots per column-marching block.  Five when a launch's shots fill blocks
// of five but not of four (the 10-shot passes of a 540 x 1040 grid): no idle
// waves, fewer accumulator sets, the model's rows read once for more shots;
// 2-3 % faster there (in-process A/B).  Four otherwise: order 8 in blocks of
// five was 25 % slower (its adjoint kernels' registers leave room for few
// resident blocks), and passes of 1-3 shots 6 % slower.
int march_spb(int sdo, int nshots)
{
    return sdo == 2 && nshots % 5 == 0 && nshots % 4 != 0 ? 5 : 4;
}

// 1-D grid of (column block, row strip, shot group) tiles for blocks of spb
// shots, XCD-interleaved (see tile()).
static dim3 march_grid(DevGrid &g, int sdo, int nshots, int spb, int span = 64,
                       bool tall_ok = false)
{
    const int nx_act = g.nx - 2 * sdo + 1, nz_act = g.nz - 2 * sdo + 1;
    g.ncb = (nx_act + span - 1) / span;
    g.nbatch = (nshots + spb - 1) / spb;
    g.nshot = nshots;
    const int rows = march_rows(g, nz_act, g.ncb, nshots, tall_ok);
    g.mrows = rows;
    g.nstrips = (nz_act + rows - 1) / rows;
    // one shot: tiles of spb column blocks (see tile_lane)
    const int T = spb > 1 && nshots == 1 ? (g.ncb + spb - 1) / spb * g.nstrips
                                         : g.ncb * g.nstrips * g.nbatch;
    const int per = (T + NUM_XCD - 1) / NUM_XCD;
    return dim3(per * NUM_XCD);
}

#define DISPATCH_SDO(sdo, KERNEL, ...)                                    \
    do {                                                                  \
        if ((sdo) == 2) KERNEL<2 __VA_ARGS__><<<grid, block, 0, st>>>(ARGS); \
        else KERNEL<4 __VA_ARGS__><<<grid, block, 0, st>>>(ARGS);          \
    } while (0)

// Column-marching kernels: half stencil width and shots per block (order 8
// always takes blocks of four, see march_spb).
#define DISPATCH_MARCH(sdo, spb, KERNEL, ...)                                          \
    do {                                                                               \
        if ((sdo) == 2 && (spb) == 5) KERNEL<2 __VA_ARGS__, 5><<<grid, block, 0, st>>>(ARGS); \
        else if ((sdo) == 2) KERNEL<2 __VA_ARGS__, 4><<<grid, block, 0, st>>>(ARGS);        \
        else KERNEL<4 __VA_ARGS__, 4><<<grid, block, 0, st>>>(ARGS);                        \
    } while (0)

void launch_vel(const DevGrid &g, int dir, int nshots, const pyfwi_shot *shots, int t,
                const float *in, int64_t in_stride, float *out, int64_t out_stride,
                const float *bu, const float *dxp, const float *dzp, hipStream_t st,
                const float *src)
{
    DevGrid gm = g;
    const int spb = march_spb(g.sdo, nshots);
    dim3 grid = march_grid(gm, g.sdo, nshots, spb), block(64 * spb);
#define ARGS gm, shots, t, in, in_stride, out, out_stride, bu, dxp, dzp, src
    if (dir > 0) DISPATCH_MARCH(g.sdo, spb, k_vel, , 1);
    else DISPATCH_MARCH(g.sdo, spb, k_vel, , -1);
#undef ARGS
}

void launch_str(const DevGrid &g, int dir, int nshots, const pyfwi_shot *shots, int t,
                const float *in, int64_t in_stride, float *out, int64_t out_stride,
                const float *lam, const float *mu, const float *dxp, const float *dzp,
                hipStream_t st, const float *src, float *seis)
{
    DevGrid gm = g;
    const int spb = march_spb(g.sdo, nshots);
    dim3 grid = march_grid(gm, g.sdo, nshots, spb), block(64 * spb);
#define ARGS gm, shots, t, in, in_stride, out, out_stride, lam, mu, dxp, dzp, src, seis
    if (dir > 0) DISPATCH_MARCH(g.sdo, spb, k_str, , 1);
    else DISPATCH_MARCH(g.sdo, spb, k_str, , -1);
#undef ARGS
}

// Grid of the fused kernels: column blocks of FUSED_WAVES wavefronts, each
// owning 64 - 2*SDO columns; strips as in march_grid.
static dim3 fused_grid(DevGrid &g, int sdo, int nshots)
{
    const int nx_act = g.nx - 2 * sdo + 1, nz_act = g.nz - 2 * sdo + 1;
    const int span = FUSED_WAVES * (64 - 2 * sdo);
    g.ncb = (nx_act + span - 1) / span;
    g.nbatch = nshots;  // fused tiles: one shot per block
    g.nshot = nshots;
    const int rows = strip_rows(g, nz_act, (int64_t)g.ncb * FUSED_WAVES * nshots,
                                MARCH_WAVES_PER_CU, MARCH_ROWS_MIN, MARCH_ROWS_MAX);
    g.mrows = rows;
    g.nstrips = (nz_act + rows - 1) / rows;
    const int T = g.ncb * g.nstrips * g.nbatch;
    const int per = (T + NUM_XCD - 1) / NUM_XCD;
    return dim3(per * NUM_XCD);
}

void launch_fused_step(const DevGrid &g, int dir, int nshots, const pyfwi_shot *shots, int t,
                       const float *in, int64_t in_stride, float *out, int64_t out_stride,
                       const float *bu, const float *lam, const float *mu, const float *dxp,
                       const float *dzp, hipStream_t st, const float *src, float *seis)
{
    DevGrid gm = g;
    if (dir > 0) {
        const int spb = march_spb(g.sdo, nshots);
        dim3 grid = march_grid(gm, g.sdo, nshots, spb, 64 - 2 * g.sdo), block(64 * spb);
#define ARGS gm, shots, t, in, in_stride, out, out_stride, bu, lam, mu, dxp, dzp, src, seis
        DISPATCH_MARCH(g.sdo, spb, k_fwd_step, );
#undef ARGS
        return;
    }
    dim3 grid = fused_grid(gm, g.sdo, nshots), block(64 * FUSED_WAVES);
#define ARGS gm, shots, t, in, out, bu, lam, mu, dxp, dzp
    DISPATCH_SDO(g.sdo, k_fused_rev, );
#undef ARGS
}

bool launch_fwd_two(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                    const float *in, int64_t in_stride, float *out0, float *out1,
                    int64_t out_stride, const float *bu, const float *lam, const float *mu,
                    const float *dxp, const float *dzp, hipStream_t st, const float *src,
                    float *seis)
{
    if (g.sdo != 2) return false;  // order 4 only
    DevGrid gm = g;
    if (gm.rows_override == 0) gm.rows_override = FWD_TWO_ROWS;
    const int spb = march_spb(g.sdo, nshots);
    dim3 grid = march_grid(gm, g.sdo, nshots, spb, 64 - 6 * g.sdo), block(64 * spb);
#define ARGS gm, shots, t, in, in_stride, out0, out1, out_stride, bu, lam, mu, dxp, dzp, src, seis
    if (spb == 5) k_fwd_two<5><<<grid, block, 0, st>>>(ARGS);
    else k_fwd_two<4><<<grid, block, 0, st>>>(ARGS);
#undef ARGS
    return true;
}

void launch_inject_record(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                          float *fields, int64_t fields_stride, const float *src, float *seis,
                          const int32_t *rtab, hipStream_t st)
{
    hipLaunchKernelGGL(k_inject_record, dim3(nshots), dim3(256), 0, st, g, shots, t, fields,
                       fields_stride, src, seis, rtab);
}

void launch_record_kept(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t0, int t1,
                        const float *kept, int64_t shot_stride, int64_t slot_stride, float *seis,
                        const int32_t *rtab, hipStream_t st)
{
    if (t1 <= t0 || nshots <= 0) return;
    // grid.y holds at most 65535 shots a launch
    for (int s0 = 0; s0 < nshots; s0 += 65535)
        hipLaunchKernelGGL(k_record_kept, dim3(t1 - t0, std::min(65535, nshots - s0)), dim3(256),
                           0, st, g, shots + s0, t0, kept + (int64_t)s0 * shot_stride, shot_stride,
                           slot_stride, seis, rtab);
}

void launch_adj_inject(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                       float *adj, const float *res, const int32_t *rtab, hipStream_t st, int f0,
                       int f1, bool first_only)
{
    hipLaunchKernelGGL(k_adj_inject, dim3(nshots), dim3(256), 0, st, g, shots, t, adj, res, f0,
                       f1, first_only ? 1 : 0, rtab);
}

void launch_adj_str(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                    const float *fields, int64_t fields_stride, float *adj, const float *bu,
                    const float *dxp, const float *dzp, float *gacc, bool precond,
                    hipStream_t st, const float *res)
{
    DevGrid gm = g;
    const int spb = march_spb(g.sdo, nshots);
    dim3 grid = march_grid(gm, g.sdo, nshots, spb), block(64 * spb);
    const bool alone = nshots == 1;  // no other shot to share the accumulators with
#define ARGS gm, shots, t, fields, fields_stride, adj, bu, dxp, dzp, gacc, res
    if (precond && alone) DISPATCH_MARCH(g.sdo, spb, k_adj_str, , true, true);
    else if (precond) DISPATCH_MARCH(g.sdo, spb, k_adj_str, , true, false);
    else if (alone) DISPATCH_MARCH(g.sdo, spb, k_adj_str, , false, true);
    else DISPATCH_MARCH(g.sdo, spb, k_adj_str, , false, false);
#undef ARGS
}

void launch_adj_vel_grad(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                         const float *fields, int64_t fields_stride, float *adj,
                         const float *lam, const float *mu, const float *bu,
                         const float *dxp, const float *dzp, float *gacc, bool precond,
                         hipStream_t st, const float *res)
{
    DevGrid gm = g;
    const int spb = march_spb(g.sdo, nshots);
    dim3 grid = march_grid(gm, g.sdo, nshots, spb), block(64 * spb);
    const bool alone = nshots == 1;
#define ARGS gm, shots, t, fields, fields_stride, adj, lam, mu, bu, dxp, dzp, gacc, res
    if (precond && alone) DISPATCH_MARCH(g.sdo, spb, k_adj_vel_grad, , true, true);
    else if (precond) DISPATCH_MARCH(g.sdo, spb, k_adj_vel_grad, , true, false);
    else if (alone) DISPATCH_MARCH(g.sdo, spb, k_adj_vel_grad, , false, true);
    else DISPATCH_MARCH(g.sdo, spb, k_adj_vel_grad, , false, false);
#undef ARGS
}

bool launch_adj_fused(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                      const float *fields, int64_t fields_stride, const float *ain, float *aout,
                      const float *lam, const float *mu, const float *bu, const float *dxp,
                      const float *dzp, float *gacc, bool precond, hipStream_t st,
                      const float *res)
{
    if (g.sdo != 2) return false;  // order 4 only
    DevGrid gm = g;
    const int spb = march_spb(g.sdo, nshots);
    dim3 grid = march_grid(gm, g.sdo, nshots, spb, 64 - 2 * g.sdo), block(64 * spb);
    const bool alone = nshots == 1;
#define ARGS gm, shots, t, fields, fields_stride, ain, aout, lam, mu, bu, dxp, dzp, gacc, res
#define LAUNCH(P, A)                                                                   \
    do {                                                                               \
        if (spb == 5) k_adj_fused<2, P, A, 5><<<grid, block, 0, st>>>(ARGS);            \
        else k_adj_fused<2, P, A, 4><<<grid, block, 0, st>>>(ARGS);                     \
    } while (0)
    if (precond && alone) LAUNCH(true, true);
    else if (precond) LAUNCH(true, false);
    else if (alone) LAUNCH(false, true);
    else LAUNCH(false, false);
#undef LAUNCH
#undef ARGS
    return true;
}

void launch_exact_str(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                      const float *vel, int64_t vel_stride, float *adj, float *av,
                      const float *lam, const float *mu, const float *dxp,
                      const float *dzp, float *gacc, hipStream_t st)
{
    DevGrid gm = g;
    const int spb = march_spb(g.sdo, nshots);
    dim3 grid = march_grid(gm, g.sdo, nshots, spb, 64, true), block(64 * spb);
    const bool alone = nshots == 1;
#define ARGS gm, shots, t, vel, vel_stride, adj, av, lam, mu, dxp, dzp, gacc
    if (alone) DISPATCH_MARCH(g.sdo, spb, k_exact_str, , true);
    else DISPATCH_MARCH(g.sdo, spb, k_exact_str, , false);
#undef ARGS
}

void launch_exact_vel(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                      const float *tau, int64_t tau_stride, float *adj, const float *av_in,
                      float *av_out, const float *bu, const float *dxp, const float *dzp,
                      float *gacc, hipStream_t st, const float *src)
{
    DevGrid gm = g;
    const int spb = march_spb(g.sdo, nshots);
    dim3 grid = march_grid(gm, g.sdo, nshots, spb, 64, true), block(64 * spb);
    const bool alone = nshots == 1;
#define ARGS gm, shots, t, tau, tau_stride, adj, av_in, av_out, bu, dxp, dzp, gacc, src
    if (alone) DISPATCH_MARCH(g.sdo, spb, k_exact_vel, , true);
    else DISPATCH_MARCH(g.sdo, spb, k_exact_vel, , false);
#undef ARGS
}

void launch_exact_inject(const DevGrid &g, int nshots, const pyfwi_shot *shots, int t,
                         float *adj, float *av, const float *res, const int32_t *rtab,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_exact_inject, dim3(nshots), dim3(256), 0, st, g, shots, t, adj,
                       av, res, rtab);
}

void launch_zero(float *p, int64_t n, hipStream_t st)
{
    if (n <= 0) return;
    if (reinterpret_cast<uintptr_t>(p) % 16) {  // a caller's buffer off the 16-byte grid
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_zero1, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
        return;
    }
    const int64_t blocks = std::min<int64_t>((n / 4 + 255) / 256 + 1, 8192);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
}

void launch_copy_shots(float *dst, int64_t dst_stride, const float *src, int64_t src_stride,
                       int64_t count, int nshots, hipStream_t st)
{
    int64_t n4 = count / 4;
    int blocks = (int)((n4 + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy_shots, dim3(blocks, nshots), dim3(256), 0, st, dst, dst_stride,
                       src, src_stride, count);
}

void launch_grad_sum(const DevGrid &g, int nsets, int nslot, const float *gacc, float *out,
                     const float *bu, int b_to_rho, int accumulate, hipStream_t st)
{
    hipLaunchKernelGGL(k_grad_sum, dim3(1024), dim3(256), 0, st, g, nsets, nslot, gacc, out,
                       bu, b_to_rho, accumulate);
}

void launch_pad(const DevGrid &g, const float *src, float *dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_pad, dim3(512), dim3(256), 0, st, g, src, dst);
}

void launch_model_check(const DevGrid &g, const float *lam, const float *mu, const float *bu,
                        int *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_model_check, dim3(512), dim3(256), 0, st, g, lam, mu, bu, out);
}

void launch_unpad(const DevGrid &g, const float *src, float *dst, int nf, hipStream_t st)
{
    hipLaunchKernelGGL(k_unpad, dim3(512), dim3(256), 0, st, g, src, dst, nf);
}

} // namespace pyfwi
