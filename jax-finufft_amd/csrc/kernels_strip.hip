// Atomic-free 2D spreading ("strip" method), the default type-1 spreader in 2D.
//
// Replaces the 2D type-1 spreading cuFINUFFT runs inside execute (called at
// lib/kernels.cc.cu:75; source absent from the reference).  Designed around
// two measured facts of gfx950: LDS float atomics are ~35x slower than LDS
// integer atomics or plain LDS traffic, and coalesced plain stores are far
// cheaper than global float atomics.
//
// Layout.  The fine grid is cut into vertical strips of SW = 64 - ns + 1
// columns; strip s owns points whose first x-tap l0x lies in
// [s*SW, (s+1)*SW), so all of their taps fall in the 64 columns
// [s*SW, s*SW + 64).  Points are sorted by (strip, l0y) (the "y-major" bins
// set up by the plan; for ns <= 8 also by 8-column block within the row), so
// within a strip they arrive in row order.  A strip is cut into units of H
// consecutive l0y rows, and one wave spreads one unit.
//
// ns <= 8 (strip8_piece): cells are assigned to lanes modulo 8 in both
// directions, so that every lane does useful work for every point (below).
// Wider kernels (wide_piece): lane x owns column s*SW + x and holds a rolling
// window of ns rows (real and imaginary parts) in registers.  Points are staged
// 64 at a time: each lane evaluates one point's ns x-weights and ns y-weights
// (Horner) and writes the x-weights and c*y-weights to wave-private LDS; then
// the wave walks the batch in order, every lane reading its own x-weight (zero
// outside the point's window) and the point's y-weights by broadcast, and doing
// 2*ns FMAs into its window.  When the next point starts a later row, the
// window's first row is complete and is written out.  No atomics touch LDS.
//
// Output.  A unit's points are cut into pieces of at most `pmax` points (an
// empty unit has none), one wave per piece.  A piece writes its rows with plain
// coalesced stores into its own scratch tile of (H + ns - 1) x 64 cells; a
// merge kernel then writes every fine-grid cell once, as the sum over the
// units covering it (one to four, up to nine next to a narrow last strip or
// unit) of their pieces' tiles.  The fine grid is
// neither zeroed first nor read, no float atomics are used anywhere, and the
// summation order is fixed, so results are bitwise reproducible.
//
// The same per-piece and per-cell bodies serve one plan (a single problem,
// blockIdx.y = transform) and a ragged batch of problems (descriptor tables,
// one launch over every problem's units).
//
// 3D.  A 3D grid is a stack of nz = nf2 slices of nf0 x nf1 cells.  A point
// with first taps (l0x, l0y, l0z) and z-weights kz[dz] spreads, in slice
// l0z + dz, like a 2D point at (l0x, l0y) of strength c * kz[dz]; setpts
// expands every point into its ns slice points and sorts them by (slice,
// strip, row), so the strips of all slices form one list of nz * S strips
// (strip index z * S + s) and the 2D bodies run unchanged.  Only the merge
// needs the slice: its neighbours are taken within the slice.
#include "jfa_device.hpp"

namespace jfa {
namespace {

constexpr int kStripLanes = 64;

template <typename T>
struct alignas(2 * sizeof(T)) Pair {
    T x, y;
};
template <typename T>
struct alignas(4 * sizeof(T)) Quad {
    T x, y, z, w;
};

// Pairs per staged point in the ns <= 8 layout: 8 and one pad.  Each lane
// writes its own point's pairs, so the point stride is the lane stride of the
// staging writes; 9 pairs (18 or 36 dwords, twice an odd number) put the
// lanes of one write on distinct banks, where 8 pairs would put every fourth
// (64 banks) or second (32 banks) lane on the same bank.
constexpr int kPointPairs = 9;

// LDS staging of one wave: ns <= 8 uses the weight pairs (and c * y-weights
// for NT transforms at once), wider kernels the column weights and c *
// y-weights per point.
template <typename T, int NS, int NT = 1, bool MOD8 = (NS <= 8)>
struct StripLds;
template <typename T, int NS, int NT>
struct StripLds<T, NS, NT, true> {
    // per point q, lane column a: (weight in block bx, in bx + 1)
    Pair<T> w[kStripLanes * kPointPairs];
    // per transform, point q, row offset dy: c * ky[dy]
    Pair<T> cky[NT][kStripLanes * kPointPairs];
};
// ns > 8: the point stride is rounded up to an odd number of values for the
// same reason.
template <int NS>
constexpr int wide_stride() { return NS | 1; }
template <typename T, int NS, int NT>
struct StripLds<T, NS, NT, false> {
    T kx[kStripLanes * wide_stride<NS>()];
    T cky[kStripLanes * wide_stride<NS>() * 2];
    int lx[kStripLanes];
    int ly[kStripLanes];
};

// Strength and Horner pair of input point P: from the packed record, or from
// the strengths (by index within the problem) and the setpts pairs (by P).
template <typename T>
__device__ __forceinline__ Quad<T> point_record(const StripPts<T> &pts, int64_t P) {
    const int64_t j = P - pts.idx_off;
    if (pts.pack) return reinterpret_cast<const Quad<T> *>(pts.pack)[j];
    const Pair<T> cv = reinterpret_cast<const Pair<T> *>(pts.c)[j];
    const Pair<T> tv = reinterpret_cast<const Pair<T> *>(pts.rec)[P];
    return Quad<T>{cv.x, cv.y, tv.x, tv.y};
}

// The point range [pb, pe) of piece `sub` of unit u, and its strip / rows.
struct PieceRange {
    int ox, r0, pb, pe;
};
__device__ __forceinline__ PieceRange piece_range(const StripGeom &sg,
                                                  const int *__restrict__ bin_start, int u,
                                                  int sub) {
    const int s = u / sg.K, k = u - s * sg.K;
    PieceRange pr;
    pr.ox = s * sg.SW;
    pr.r0 = k * sg.H;
    const int r1 = min(pr.r0 + sg.H, (int)sg.nf1);
    const int ub = bin_start[(int64_t)s * sg.nf1 + pr.r0];
    const int ue = bin_start[(int64_t)s * sg.nf1 + r1];
    pr.pb = ub + sub * sg.pmax;
    pr.pe = min(pr.pb + sg.pmax, ue);
    return pr;
}

// ------------------------------------------------------------ ns > 8 body

template <typename T, int NS, typename Lds>
__device__ __forceinline__ void wide_piece(const StripGeom &sg, const StripPts<T> &pts,
                                           const int *__restrict__ bin_start, int u, int sub,
                                           const T *__restrict__ coef, T beta, int horner,
                                           T *__restrict__ tile, Lds &lds) {
    constexpr int NSP = wide_stride<NS>();
    const int lane = threadIdx.x;
    const int rows_tile = sg.H + NS - 1;
    const PieceRange pr = piece_range(sg, bin_start, u, sub);
    const int r0 = pr.r0, pb = pr.pb, pe = pr.pe;
    // key >> kStripKeyShift = rowbase + l0y for every point of this strip
    const int64_t rowbase = pts.bin_off + (int64_t)(u / sg.K) * sg.nf1;
    T acc_r[NS], acc_i[NS];
#pragma unroll
    for (int b = 0; b < NS; ++b) acc_r[b] = acc_i[b] = T(0);
    // rolling window: acc[b] holds row rcur + b
    int rcur = r0;
    auto flush = [&](int row) {
        {
            const int64_t q = (int64_t)(row - r0) * kStripLanes + lane;
            tile[2 * q] = acc_r[0];
            tile[2 * q + 1] = acc_i[0];
        }
#pragma unroll
        for (int b = 0; b + 1 < NS; ++b) {
            acc_r[b] = acc_r[b + 1];
            acc_i[b] = acc_i[b + 1];
        }
        acc_r[NS - 1] = acc_i[NS - 1] = T(0);
    };
    for (int bb = pb; bb < pe; bb += kStripLanes) {
        const int nb = min(kStripLanes, pe - bb);
        // stage: lane q evaluates point bb + q
        if (lane < nb) {
            const int p = bb + lane;
            const unsigned key = pts.key[p];
            const Quad<T> rq = point_record(pts, pts.idx[p]);
            const T cr = rq.x, ci = rq.y;
            const Pair<T> tt{rq.z, rq.w};
            T kx[NS], ky[NS];
            if (horner) {
                eval_kernel<T, NS, true>(coef, beta, tt.x, kx);
                eval_kernel<T, NS, true>(coef, beta, tt.y, ky);
            } else {
                eval_kernel<T, NS, false>(coef, beta, tt.x, kx);
                eval_kernel<T, NS, false>(coef, beta, tt.y, ky);
            }
#pragma unroll
            for (int a = 0; a < NS; ++a) lds.kx[lane * NSP + a] = kx[a];
#pragma unroll
            for (int b = 0; b < NS; ++b) {
                lds.cky[(lane * NSP + b) * 2] = cr * ky[b];
                lds.cky[(lane * NSP + b) * 2 + 1] = ci * ky[b];
            }
            lds.lx[lane] = (int)(key & ((1u << kStripKeyShift) - 1));
            lds.ly[lane] = (int)((key >> kStripKeyShift) - rowbase);
        }
        __syncthreads();
        for (int q = 0; q < nb; ++q) {
            const int ly = lds.ly[q];
            while (rcur < ly) {
                flush(rcur);
                ++rcur;
            }
            const int kk = lane - lds.lx[q];
            const bool in = kk >= 0 && kk < NS;
            const T kxv = in ? lds.kx[q * NSP + (in ? kk : 0)] : T(0);
            const T *cy = &lds.cky[q * NSP * 2];
#pragma unroll
            for (int b = 0; b < NS; ++b) {
                acc_r[b] = fma(cy[2 * b], kxv, acc_r[b]);
                acc_i[b] = fma(cy[2 * b + 1], kxv, acc_i[b]);
            }
        }
        __syncthreads();
    }
    // drain the window and every remaining row of the tile
    const int rstop = r0 + rows_tile;
    while (rcur < rstop) {
        flush(rcur);
        ++rcur;
    }
}

// ------------------------------------------------------------ ns <= 8 body
//
// Points are staged 64 at a time in wave-private LDS: per point q, its
// x-weights placed in a 16-column frame starting at the 8-column block
// bx = X >> 3 of its first tap X (frame column c = (X & 7) + i holds kx[i];
// stored as pairs by c & 7: (weight at c, weight at c + 8)), and its
// strength times its y-weights, c * ky[dy].  The wave then walks the batch in
// runs of equal (row, bx): the sort key's low bits order the points by row and
// block, so every point of a run adds into the same two column blocks of an
// 8-row window of 64 columns held in registers (row y in window slot y & 7).
// When a later row starts, the window's first row is complete: it is written
// to the piece's tile and its slot cleared.
//
// Lane (a, r) = (lane & 7, lane >> 3) holds columns a, a+8, ..., a+56 of the
// row in slot r.  An ns x ns footprint (ns <= 8) puts at most one cell in
// every lane, so per point each lane adds the products of its weight pair
// with the point's c * ky for its row into blocks bx and bx + 1.

// The accumulator window (see above).
template <typename T, int NT>
struct Window8 {
    T re[NT][8], im[NT][8];

    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int b = 0; b < 8; ++b) re[t][b] = im[t][b] = T(0);
    }

    // points [q0, q1) of the staged batch, all in row Y and block BX
    template <int BX>
    __device__ __forceinline__ void walk(const Pair<T> *__restrict__ s_w,
                                         const Pair<T> (*__restrict__ s_cky)
                                             [kStripLanes * kPointPairs],
                                         int q0, int q1, int Y, int lane) {
        const int a = lane & 7, dy = ((lane >> 3) - Y) & 7;
#pragma unroll 4
        for (int q = q0; q < q1; ++q) {
            const Pair<T> w = s_w[q * kPointPairs + a];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const Pair<T> cy = s_cky[t][q * kPointPairs + dy];
                re[t][BX] = fma(w.x, cy.x, re[t][BX]);
                im[t][BX] = fma(w.x, cy.y, im[t][BX]);
                if constexpr (BX < 7) {  // block 8 would only receive zero weights
                    re[t][BX + 1] = fma(w.y, cy.x, re[t][BX + 1]);
                    im[t][BX + 1] = fma(w.y, cy.y, im[t][BX + 1]);
                }
            }
        }
    }

    // write window row `row` (tile row row - r0) and clear its slot
    __device__ __forceinline__ void flush(int row, int r0, int lane, T *__restrict__ tile,
                                          int64_t tile_tstride) {
        if ((row & 7) != (lane >> 3)) return;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            Pair<T> *dst = reinterpret_cast<Pair<T> *>(tile + t * tile_tstride) +
                           (int64_t)(row - r0) * kStripLanes + (lane & 7);
#pragma unroll
            for (int b = 0; b < 8; ++b) dst[8 * b] = Pair<T>{re[t][b], im[t][b]};
#pragma unroll
            for (int b = 0; b < 8; ++b) re[t][b] = im[t][b] = T(0);
        }
    }
};

// One piece for NT transforms sharing the points: transform k's strengths at
// pts.c + k * pts.c_tstride (NT = 1 may read packed records instead), its tile
// at tile + k * tile_tstride.
template <typename T, int NS, int NT, int LNT>
__device__ __forceinline__ void strip8_piece(const StripGeom &sg, const StripPts<T> &pts,
                                             const int *__restrict__ bin_start, int u, int sub,
                                             const T *__restrict__ coef, T beta, int horner,
                                             T *__restrict__ tile, int64_t tile_tstride,
                                             StripLds<T, NS, LNT> &lds) {
    static_assert(NS <= 8, "modular ownership needs ns <= 8");
    static_assert(NT <= LNT, "LDS staging for NT transforms");
    Pair<T> *s_w = lds.w;
    Pair<T> (*s_cky)[kStripLanes * kPointPairs] = lds.cky;
    const int lane = threadIdx.x;
    const int rows_tile = sg.H + NS - 1;
    const PieceRange pr = piece_range(sg, bin_start, u, sub);
    const int r0 = pr.r0, pb = pr.pb, pe = pr.pe;
    // key >> kStripKeyShift = rowbase + l0y for every point of this strip
    const int64_t rowbase = pts.bin_off + (int64_t)(u / sg.K) * sg.nf1;
    Window8<T, NT> win;
    win.clear();
    // the window holds rows R .. R+7
    int R = r0;
    // Software pipeline over 64-point batches: while one batch is staged and
    // walked, the next batch's records (and other transforms' strengths) and
    // the key and index of the batch after it are in flight.  A wave holds one
    // unit, so without this every batch would wait for its scattered record
    // reads with too few other waves on the SIMD to cover them.
    constexpr int NC = NT > 1 ? NT - 1 : 1;
    unsigned key1 = 0, key2 = 0;
    int idx1 = 0, idx2 = 0;
    Quad<T> rq1{};
    Pair<T> ct1[NC];
    auto load_index = [&](int b, unsigned &k, int &i) {
        if (b + lane < pe) {
            k = pts.key[b + lane];
            i = pts.idx[b + lane];
        }
    };
    auto load_record = [&](int b, int i) {
        if (b + lane < pe) {
            rq1 = point_record(pts, i);
#pragma unroll
            for (int t = 1; t < NT; ++t)
                ct1[t - 1] = reinterpret_cast<const Pair<T> *>(
                    pts.c + t * pts.c_tstride)[i - pts.idx_off];
        }
    };
    load_index(pb, key1, idx1);
    load_record(pb, idx1);
    load_index(pb + kStripLanes, key2, idx2);
    for (int bb = pb; bb < pe; bb += kStripLanes) {
        const int nb = min(kStripLanes, pe - bb);
        const unsigned pkey = key1;
        const Quad<T> rq = rq1;
        Pair<T> ct[NC];
#pragma unroll
        for (int t = 0; t < NC; ++t) ct[t] = ct1[t];
        key1 = key2;
        idx1 = idx2;
        load_record(bb + kStripLanes, idx1);
        load_index(bb + 2 * kStripLanes, key2, idx2);
        int key = -1;  // (row << 3) | block of the lane's staged point
        if (lane < nb) {
            const T cr = rq.x, ci = rq.y;
            const Pair<T> tt{rq.z, rq.w};
            const int X = (int)(pkey & ((1u << kStripKeyShift) - 1));
            // x-weights, then y-weights: one set of ns values live at a time
            {
                T kx[NS];
                if (horner)
                    eval_kernel<T, NS, true>(coef, beta, tt.x, kx);
                else
                    eval_kernel<T, NS, false>(coef, beta, tt.x, kx);
                const int ph = X & 7;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const T v = i < NS ? kx[i] : T(0);
                    const bool lo = ph + i < 8;
                    s_w[lane * kPointPairs + ((ph + i) & 7)] =
                        Pair<T>{lo ? v : T(0), lo ? T(0) : v};
                }
            }
            {
                T ky[NS];
                if (horner)
                    eval_kernel<T, NS, true>(coef, beta, tt.y, ky);
                else
                    eval_kernel<T, NS, false>(coef, beta, tt.y, ky);
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    s_cky[0][lane * kPointPairs + i] =
                        i < NS ? Pair<T>{cr * ky[i], ci * ky[i]} : Pair<T>{T(0), T(0)};
#pragma unroll
                for (int t = 1; t < NT; ++t) {
#pragma unroll
                    for (int i = 0; i < 8; ++i)
                        s_cky[t][lane * kPointPairs + i] =
                            i < NS ? Pair<T>{ct[t - 1].x * ky[i], ct[t - 1].y * ky[i]}
                                   : Pair<T>{T(0), T(0)};
                }
            }
            key = ((int)((pkey >> kStripKeyShift) - rowbase) << 3) | (X >> 3);
        }
        const int prev = __shfl_up(key, 1);
        const uint64_t starts = __ballot(lane < nb && (lane == 0 || key != prev));
        __syncthreads();
        uint64_t mask = starts;
        while (mask) {
            const int q0 = __builtin_ctzll(mask);
            mask &= mask - 1;
            const int q1 = mask ? __builtin_ctzll(mask) : nb;
            const int gkey = __builtin_amdgcn_readlane(key, q0);
            const int Y = gkey >> 3;
            while (R < Y) {
                win.flush(R, r0, lane, tile, tile_tstride);
                ++R;
            }
            switch (gkey & 7) {
                case 0: win.template walk<0>(s_w, s_cky, q0, q1, Y, lane); break;
                case 1: win.template walk<1>(s_w, s_cky, q0, q1, Y, lane); break;
                case 2: win.template walk<2>(s_w, s_cky, q0, q1, Y, lane); break;
                case 3: win.template walk<3>(s_w, s_cky, q0, q1, Y, lane); break;
                case 4: win.template walk<4>(s_w, s_cky, q0, q1, Y, lane); break;
                case 5: win.template walk<5>(s_w, s_cky, q0, q1, Y, lane); break;
                case 6: win.template walk<6>(s_w, s_cky, q0, q1, Y, lane); break;
                default: win.template walk<7>(s_w, s_cky, q0, q1, Y, lane); break;
            }
        }
        __syncthreads();
    }
    // drain the window and every remaining row of the tile
    const int rstop = r0 + rows_tile;
    while (R < rstop) {
        win.flush(R, r0, lane, tile, tile_tstride);
        ++R;
    }
}

// One piece for nt transforms (nt <= LNT) sharing the points (see strip8_piece).
template <typename T, int NS, int LNT>
__device__ __forceinline__ void strip_piece(const StripGeom &sg, const StripPts<T> &pts,
                                            const int *__restrict__ bin_start, int u, int sub,
                                            const T *__restrict__ coef, T beta, int horner,
                                            T *__restrict__ tile, int nt, int64_t tile_tstride,
                                            StripLds<T, NS, LNT> &lds) {
    if constexpr (NS <= 8) {
        if (LNT > 1 && nt > 1)
            strip8_piece<T, NS, LNT, LNT>(sg, pts, bin_start, u, sub, coef, beta, horner, tile,
                                          tile_tstride, lds);
        else
            strip8_piece<T, NS, 1, LNT>(sg, pts, bin_start, u, sub, coef, beta, horner, tile,
                                        tile_tstride, lds);
    } else {
        StripPts<T> pk = pts;
        for (int k = 0; k < nt; ++k) {
            wide_piece<T, NS>(sg, pk, bin_start, u, sub, coef, beta, horner,
                              tile + k * tile_tstride, lds);
            pk.c += pts.c_tstride;
        }
    }
}

// The windows of one axis that reach into window w's first ns - 1 cells from
// the left: window w - 1, or, for w = 0, the last one or two windows through
// the periodic wrap (the second when the last window owns fewer than ns - 1
// cells).  Windows start every `step` cells, span `span` and n of them cover
// nf cells.  Appends (window, offset of w's cell 0 in its tile) to win / off.
__device__ __forceinline__ int left_windows(int w, int step, int span, int n, int64_t nf,
                                            int *win, int *off) {
    int c = 0;
    if (w > 0) {
        win[c] = w - 1;
        off[c] = step;
        ++c;
    } else {
        for (int v = n - 1; v >= 1 && v >= n - 2; --v) {
            const int o = (int)(nf - (int64_t)v * step);
            if (o < span) {
                win[c] = v;
                off[c] = o;
                ++c;
            }
        }
    }
    return c;
}

// The cells unit u owns (columns [s*SW, s*SW + SW) and rows [k*H, k*H + H),
// clipped to the grid), each written once as the sum of the tiles of every
// piece of every unit covering it: u itself, and the units to its left, above
// and diagonally where their tiles reach into u's first ns - 1 columns / rows.
// One workgroup per unit, lane = column, waves over rows; the candidate units
// and their piece ranges are uniform over the workgroup.  piece_start holds the
// problem's units' first piece numbers; tile i of `scratch` belongs to piece
// piece_start[0] + i.
template <typename T, int NS>
__device__ __forceinline__ void merge_unit(const StripGeom &sg, const int *__restrict__ piece_start,
                                           const T *__restrict__ scratch, T *__restrict__ fw,
                                           int u) {
    constexpr int SW = 65 - NS;
    const int sz = u / sg.K, k = u - sz * sg.K;  // strip over all slices, unit row
    const int z = sz / sg.S, s = sz - z * sg.S;  // slice, strip within it
    fw += 2 * (int64_t)z * sg.nf0 * sg.nf1;
    const int x0 = s * SW, y0 = k * sg.H;
    const int wcols = min(SW, (int)(sg.nf0 - x0));
    const int wrows = min(sg.H, (int)(sg.nf1 - y0));
    const int rows_tile = sg.H + NS - 1;
    const int64_t tile_cells = (int64_t)rows_tile * kStripLanes;
    const int base = piece_start[0];
    int cs[3] = {s, 0, 0}, co[3] = {0, 0, 0}, ck[3] = {k, 0, 0}, ro[3] = {0, 0, 0};
    const int ncs = 1 + left_windows(s, SW, kStripLanes, sg.S, sg.nf0, cs + 1, co + 1);
    const int nck = 1 + left_windows(k, sg.H, rows_tile, sg.K, sg.nf1, ck + 1, ro + 1);
    // the covering units' piece ranges, once per workgroup (uniform); entry
    // 3 i + j is strip candidate i with unit-row candidate j (no pieces when
    // either does not exist)
    int pfirst[9], pcount[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            pfirst[3 * i + j] = 0;
            pcount[3 * i + j] = 0;
            if (i < ncs && j < nck) {
                const int u2 = (z * sg.S + cs[i]) * sg.K + ck[j];
                pfirst[3 * i + j] = piece_start[u2] - base;
                // a crowded unit's pieces were summed into its first tile
                pcount[3 * i + j] = min(piece_start[u2 + 1] - piece_start[u2], 1);
            }
        }
    const int c = threadIdx.x & (kStripLanes - 1);
    if (c >= wcols) return;
    const Pair<T> *tiles = reinterpret_cast<const Pair<T> *>(scratch);
    for (int r = threadIdx.x / kStripLanes; r < wrows; r += blockDim.x / kStripLanes) {
        T ar = T(0), ai = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int q = 3 * i + j;
                const int tc = c + co[i], tr = r + ro[j];
                if (pcount[q] == 0 || tc >= kStripLanes || tr >= rows_tile) continue;
                const Pair<T> *src =
                    tiles + (int64_t)pfirst[q] * tile_cells + tr * kStripLanes + tc;
                for (int pc = 0; pc < pcount[q]; ++pc) {
                    const Pair<T> v = src[pc * tile_cells];
                    ar += v.x;
                    ai += v.y;
                }
            }
        reinterpret_cast<Pair<T> *>(fw)[(int64_t)(y0 + r) * sg.nf0 + x0 + c] = Pair<T>{ar, ai};
    }
}

// pieces of unit u: ceil(count / pmax), none for an empty unit
__device__ __forceinline__ int unit_pieces(const StripGeom &sg, const int *__restrict__ bin_start,
                                           int u) {
    const int s = u / sg.K, k = u - s * sg.K;
    const int r0 = k * sg.H, r1 = min(r0 + sg.H, (int)sg.nf1);
    const int cnt = bin_start[(int64_t)s * sg.nf1 + r1] - bin_start[(int64_t)s * sg.nf1 + r0];
    return (cnt + sg.pmax - 1) / sg.pmax;
}

// One workgroup: exclusive prefix of pieces(u) over u in [0, n); total at
// piece_start[n] and in *npieces; with piece_unit, also the unit of every piece;
// with heavy, the list [count, u ...] of the units with more than one piece.
template <typename PiecesOf>
__device__ __forceinline__ void piece_scan(int n, PiecesOf pieces, int *piece_start,
                                           int *npieces, int *piece_unit, int *heavy) {
    __shared__ int sa[1024];
    const int tid = threadIdx.x;
    if (heavy && tid == 0) heavy[0] = 0;  // ordered before the atomics by the barriers below
    const int per = (n + 1023) / 1024;
    const int lo = min(tid * per, n);
    const int hi = min(lo + per, n);
    int a = 0;
    for (int u = lo; u < hi; ++u) a += pieces(u);
    sa[tid] = a;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = tid >= off ? sa[tid - off] : 0;
        __syncthreads();
        sa[tid] += v;
        __syncthreads();
    }
    int e = sa[tid] - a;
    for (int u = lo; u < hi; ++u) {
        piece_start[u] = e;
        const int np = pieces(u);
        if (piece_unit)
            for (int i = 0; i < np; ++i) piece_unit[e + i] = u;
        if (heavy && np > 1) heavy[1 + atomicAdd(&heavy[0], 1)] = u;
        e += np;
    }
    if (tid == 1023) {
        piece_start[n] = sa[1023];
        *npieces = sa[1023];
    }
}

// ------------------------------------------------------------ one plan

template <typename T, int NS>
__global__ void __launch_bounds__(kStripLanes)
    spread_strip_kernel(StripGeom sg, StripPts<T> pts, const int *__restrict__ bin_start,
                        const int *__restrict__ piece_start, const int *__restrict__ npieces,
                        const int *__restrict__ piece_unit, const T *__restrict__ coef, T beta,
                        int horner, int64_t pack_stride, T *__restrict__ scratch,
                        int64_t scratch_stride) {
    __shared__ StripLds<T, NS> lds;
    pts.pack += 4 * pack_stride * blockIdx.y;
    scratch += 2 * scratch_stride * blockIdx.y;
    const int64_t tile_cells = (int64_t)(sg.H + NS - 1) * kStripLanes;
    const int total = *npieces;
    for (int piece = blockIdx.x; piece < total; piece += gridDim.x) {
        const int u = piece_unit[piece];
        const int sub = piece - piece_start[u];
        strip_piece<T, NS>(sg, pts, bin_start, u, sub, coef, beta, horner,
                           scratch + 2 * piece * tile_cells, 1, 0, lds);
    }
}

// pack[t][j] = (c[t][j], rec[j])
template <typename T>
__global__ void __launch_bounds__(256)
    strip_pack_kernel(const T *__restrict__ c, int64_t c_stride, const T *__restrict__ rec,
                      int64_t M, T *__restrict__ pack) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= M) return;
    const Pair<T> cv = reinterpret_cast<const Pair<T> *>(c + 2 * c_stride * blockIdx.y)[j];
    const Pair<T> tv = reinterpret_cast<const Pair<T> *>(rec)[j];
    reinterpret_cast<Quad<T> *>(pack + 4 * M * blockIdx.y)[j] = Quad<T>{cv.x, cv.y, tv.x, tv.y};
}

// 3D: pack[t][j * ns + dz] = (c[t][j] * kz_j[dz], tx_j, ty_j), with the
// z-weights evaluated here from tz_j (rec3[3j .. 3j+2] = tx, ty, tz).
template <typename T, int NS>
__global__ void __launch_bounds__(256)
    strip3_pack_kernel(const T *__restrict__ c, int64_t c_stride, const T *__restrict__ rec3,
                       const T *__restrict__ coef, T beta, int horner, int64_t M,
                       T *__restrict__ pack) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= M) return;
    const Pair<T> cv = reinterpret_cast<const Pair<T> *>(c + 2 * c_stride * blockIdx.y)[j];
    const T tx = rec3[3 * j], ty = rec3[3 * j + 1], tz = rec3[3 * j + 2];
    T kz[NS];
    if (horner)
        eval_kernel<T, NS, true>(coef, beta, tz, kz);
    else
        eval_kernel<T, NS, false>(coef, beta, tz, kz);
    Quad<T> *out = reinterpret_cast<Quad<T> *>(pack + 4 * NS * M * blockIdx.y) + j * NS;
#pragma unroll
    for (int dz = 0; dz < NS; ++dz) out[dz] = Quad<T>{cv.x * kz[dz], cv.y * kz[dz], tx, ty};
}

template <typename T, int NS>
__global__ void __launch_bounds__(256)
    strip_merge_kernel(StripGeom sg, const int *__restrict__ piece_start,
                       const T *__restrict__ scratch, int64_t scratch_stride, T *__restrict__ fw,
                       int64_t fw_stride) {
    merge_unit<T, NS>(sg, piece_start, scratch + 2 * scratch_stride * blockIdx.y,
                      fw + 2 * fw_stride * blockIdx.y, (int)blockIdx.x);
}

// Piece table of one plan's units in three launches (counts with per-block
// sums, a scan of the block sums, block-local scans), for the many units of
// 3D grids; see the ragged form below.
__global__ void __launch_bounds__(256)
    strip_units_count_kernel(StripGeom sg, const int *__restrict__ bin_start, int *piece_start,
                             int *blocksum);
__global__ void __launch_bounds__(1024)
    strip_units_blocks_kernel(int nu, int nblk, const int *__restrict__ blocksum, int *blockoff,
                              int *piece_start, int *npieces, int *heavy);
__global__ void __launch_bounds__(256)
    strip_units_scan_kernel(int nu, const int *__restrict__ blockoff, int *piece_start,
                            int *piece_unit, int *heavy);

// Sum the tiles of every piece of a crowded unit into its first piece's tile
// (in piece order), so that the merge reads one tile per unit; work items
// (unit, 256-cell chunk of the tile) over a fixed grid.
template <typename T, int NS>
__global__ void __launch_bounds__(256)
    strip_reduce_kernel(StripGeom sg, const int *__restrict__ piece_start,
                        const int *__restrict__ heavy, T *__restrict__ scratch,
                        int64_t scratch_stride) {
    scratch += 2 * scratch_stride * blockIdx.y;
    const int64_t tile_cells = (int64_t)(sg.H + NS - 1) * kStripLanes;
    const int chunks = (int)((tile_cells + 255) / 256);
    const int items = heavy[0] * chunks;
    Pair<T> *tiles = reinterpret_cast<Pair<T> *>(scratch);
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        const int u = heavy[1 + w / chunks];
        const int64_t cell = (int64_t)(w % chunks) * 256 + threadIdx.x;
        if (cell >= tile_cells) continue;
        const int first = piece_start[u], cnt = piece_start[u + 1] - first;
        Pair<T> acc = tiles[first * tile_cells + cell];
        for (int pc = 1; pc < cnt; ++pc) {
            const Pair<T> v = tiles[(first + pc) * tile_cells + cell];
            acc.x += v.x;
            acc.y += v.y;
        }
        tiles[first * tile_cells + cell] = acc;
    }
}

// ------------------------------------------------------------ ragged batch
//
// Virtual problem v = (problem, transform).  Its units are the global units
// [vunit_off[v], vunit_off[v+1]); its points are its problem's, found through
// the problem's slice of the global bins; its fine grid and scratch tiles are
// slices of the concatenated buffers; its strengths come from the pointer
// table (c of problem i at ptrs[2*nprob + i], transform-major).

// Piece table of every virtual problem's units in three launches: counts and
// per-block sums, a scan of the block sums, then the block-local scans.
__device__ __forceinline__ int ragged_unit_pieces(const RaggedVirt *__restrict__ virt,
                                                  const int *__restrict__ unit_virt,
                                                  const int *__restrict__ bin_start, int gu) {
    const int v = unit_virt[gu];
    const RaggedVirt &d = virt[v];
    return unit_pieces(d.sg, bin_start + d.bin_off, gu - d.unit_off);
}

// exclusive scan of one value per thread over a 256-thread block; returns the
// block total in *total
__device__ __forceinline__ int block_scan256(int v, int *total) {
    __shared__ int sa[256];
    const int tid = threadIdx.x;
    sa[tid] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int w = tid >= off ? sa[tid - off] : 0;
        __syncthreads();
        sa[tid] += w;
        __syncthreads();
    }
    *total = sa[255];
    return sa[tid] - v;
}

__global__ void __launch_bounds__(256)
    strip_units_count_kernel(StripGeom sg, const int *__restrict__ bin_start, int *piece_start,
                             int *blocksum) {
    const int u = blockIdx.x * 256 + threadIdx.x;
    const int c = u < sg.nunits ? unit_pieces(sg, bin_start, u) : 0;
    if (u < sg.nunits) piece_start[u] = c;
    int total;
    block_scan256(c, &total);
    if (threadIdx.x == 0) blocksum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024)
    strip_units_blocks_kernel(int nu, int nblk, const int *__restrict__ blocksum, int *blockoff,
                              int *piece_start, int *npieces, int *heavy) {
    piece_scan(
        nblk, [&](int b) { return blocksum[b]; }, blockoff, npieces, nullptr, nullptr);
    if (threadIdx.x == 1023) piece_start[nu] = blockoff[nblk];
    if (threadIdx.x == 0) heavy[0] = 0;  // the scan kernel that follows appends to it
}

__global__ void __launch_bounds__(256)
    strip_units_scan_kernel(int nu, const int *__restrict__ blockoff, int *piece_start,
                            int *piece_unit, int *heavy) {
    const int u = blockIdx.x * 256 + threadIdx.x;
    const int c = u < nu ? piece_start[u] : 0;
    int total;
    const int e = block_scan256(c, &total);
    if (u < nu) {
        const int first = blockoff[blockIdx.x] + e;
        piece_start[u] = first;
        for (int i = 0; i < c; ++i) piece_unit[first + i] = u;
        if (c > 1) heavy[1 + atomicAdd(&heavy[0], 1)] = u;
    }
}

__global__ void __launch_bounds__(256)
    ragged_units_count_kernel(const RaggedVirt *__restrict__ virt,
                              const int *__restrict__ vunit_off,
                              const int *__restrict__ unit_virt, int nv,
                              const int *__restrict__ bin_start, int *piece_start,
                              int *blocksum) {
    const int nu = vunit_off[nv];
    const int gu = blockIdx.x * 256 + threadIdx.x;
    const int c = gu < nu ? ragged_unit_pieces(virt, unit_virt, bin_start, gu) : 0;
    if (gu < nu) piece_start[gu] = c;
    int total;
    block_scan256(c, &total);
    if (threadIdx.x == 0) blocksum[blockIdx.x] = total;
}

// one workgroup: blockoff = exclusive scan of the nblk block sums (nblk + 1
// entries); total also at piece_start[nu] and *npieces
__global__ void __launch_bounds__(1024)
    ragged_units_blocks_kernel(const int *__restrict__ vunit_off, int nv, int nblk,
                               const int *__restrict__ blocksum, int *blockoff,
                               int *piece_start, int *npieces, int *heavy) {
    piece_scan(
        nblk, [&](int b) { return blocksum[b]; }, blockoff, npieces, nullptr, nullptr);
    if (threadIdx.x == 0) heavy[0] = 0;  // the scan kernel that follows appends to it
    if (threadIdx.x == 1023) piece_start[vunit_off[nv]] = blockoff[nblk];
}

__global__ void __launch_bounds__(256)
    ragged_units_scan_kernel(const int *__restrict__ vunit_off, int nv,
                             const int *__restrict__ blockoff, int *piece_start,
                             int *piece_unit, int *heavy) {
    const int nu = vunit_off[nv];
    const int gu = blockIdx.x * 256 + threadIdx.x;
    const int c = gu < nu ? piece_start[gu] : 0;
    int total;
    const int e = block_scan256(c, &total);
    if (gu < nu) {
        const int first = blockoff[blockIdx.x] + e;
        piece_start[gu] = first;
        for (int i = 0; i < c; ++i) piece_unit[first + i] = gu;
        if (c > 1) heavy[1 + atomicAdd(&heavy[0], 1)] = gu;
    }
}

template <typename T, int NS>
__global__ void __launch_bounds__(kStripLanes)
    ragged_strip_kernel(const RaggedVirt *__restrict__ virt, const int *__restrict__ unit_virt,
                        StripPts<T> pts, const int *__restrict__ bin_start,
                        const int *__restrict__ piece_start, const int *__restrict__ npieces,
                        const int *__restrict__ piece_unit,
                        const T *__restrict__ coef, T beta, int horner,
                        const void *const *__restrict__ ptrs, int nprob,
                        T *__restrict__ scratch_all) {
    __shared__ StripLds<T, NS, 2> lds;
    const int total = *npieces;
    for (int piece = blockIdx.x; piece < total; piece += gridDim.x) {
        const int gu = piece_unit[piece];
        const int v = unit_virt[gu];
        const RaggedVirt d = virt[v];
        if (d.nt == 0) continue;  // spread together with transform t - 1
        const int sub = piece - piece_start[gu];
        StripPts<T> vp = pts;
        vp.c = (const T *)ptrs[2 * nprob + d.prob] + 2 * (int64_t)d.t * d.m;
        vp.c_tstride = 2 * d.m;
        vp.bin_off = d.bin_off;
        vp.idx_off = d.moff;
        const int64_t tile_cells = (int64_t)(d.sg.H + NS - 1) * kStripLanes;
        T *tile = scratch_all + 2 * (d.tile_off + (piece - piece_start[d.unit_off]) * tile_cells);
        strip_piece<T, NS>(d.sg, vp, bin_start + d.bin_off, gu - d.unit_off, sub, coef, beta,
                           horner, tile, d.nt, 2 * d.tile_next, lds);
    }
}

// Ragged form of strip_reduce_kernel: heavy lists global units.
template <typename T, int NS>
__global__ void __launch_bounds__(256)
    ragged_reduce_kernel(const RaggedVirt *__restrict__ virt, const int *__restrict__ unit_virt,
                         const int *__restrict__ piece_start, const int *__restrict__ heavy,
                         int chunks, T *__restrict__ scratch_all) {
    const int items = heavy[0] * chunks;
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        const int gu = heavy[1 + w / chunks];
        const RaggedVirt &d = virt[unit_virt[gu]];
        const int64_t tile_cells = (int64_t)(d.sg.H + NS - 1) * kStripLanes;
        const int64_t cell = (int64_t)(w % chunks) * 256 + threadIdx.x;
        if (cell >= tile_cells) continue;
        Pair<T> *tiles = reinterpret_cast<Pair<T> *>(scratch_all + 2 * d.tile_off);
        const int first = piece_start[gu] - piece_start[d.unit_off];
        const int cnt = piece_start[gu + 1] - piece_start[gu];
        Pair<T> acc = tiles[first * tile_cells + cell];
        for (int pc = 1; pc < cnt; ++pc) {
            const Pair<T> v = tiles[(first + pc) * tile_cells + cell];
            acc.x += v.x;
            acc.y += v.y;
        }
        tiles[first * tile_cells + cell] = acc;
    }
}

// block b merges global unit b (unit b - unit_off of its virtual problem)
template <typename T, int NS>
__global__ void __launch_bounds__(256)
    ragged_merge_kernel(const RaggedVirt *__restrict__ virt, const int *__restrict__ unit_virt,
                        const int *__restrict__ piece_start, const T *__restrict__ scratch_all,
                        T *__restrict__ fw_all) {
    const int gu = (int)blockIdx.x;
    const int v = unit_virt[gu];
    const RaggedVirt d = virt[v];
    merge_unit<T, NS>(d.sg, piece_start + d.unit_off, scratch_all + 2 * d.tile_off,
                      fw_all + 2 * d.fw_off, gu - d.unit_off);
}

}  // namespace

template <typename T>
int launch_spread_strip(const StripGeom &sg, const StripPts<T> &pts, const int *bin_start,
                        int *piece_start, int *npieces, int *piece_unit, int *heavy,
                        int *blocksum, const T *coef, int kerevalmeth, double beta,
                        int64_t pack_stride, T *fw,
                        int64_t fw_stride, T *scratch, int64_t scratch_stride, int ntr,
                        int phases, hipStream_t stream) {
    if (ntr == 0 || pts.M == 0) return JFA_OK;
    if (phases & kStripUnits) {
        const unsigned nblk = (unsigned)((sg.nunits + 255) / 256);
        int *blockoff = blocksum + nblk;  // nblk + 1 entries
        strip_units_count_kernel<<<nblk, 256, 0, stream>>>(sg, bin_start, piece_start, blocksum);
        strip_units_blocks_kernel<<<1, 1024, 0, stream>>>(sg.nunits, (int)nblk, blocksum,
                                                          blockoff, piece_start, npieces, heavy);
        strip_units_scan_kernel<<<nblk, 256, 0, stream>>>(sg.nunits, blockoff, piece_start,
                                                          piece_unit, heavy);
        JFA_HIP_CHECK(hipGetLastError());
    }
    // one workgroup per piece (see launch_ragged_strip); 3D slice grids have
    // tens of thousands of units
    const int64_t ub = sg.nunits + pts.M / sg.pmax + 1;
    const unsigned grid = (unsigned)(ub < (1 << 30) ? ub : (1 << 30));
    const T b = (T)beta;
    JFA_DISPATCH_NS(sg.ns, {
        if (phases & kStripSpread)
            spread_strip_kernel<T, NS_><<<dim3(grid, ntr), kStripLanes, 0, stream>>>(
                sg, pts, bin_start, piece_start, npieces, piece_unit, coef, b, kerevalmeth,
                pack_stride, scratch, scratch_stride);
        if (phases & kStripMerge) {
            const int64_t rb = strip_max_heavy(sg, pts.M) *
                               ((strip_tile_cells(sg) + 255) / 256);
            strip_reduce_kernel<T, NS_>
                <<<dim3((unsigned)(rb < 2048 ? rb : 2048), ntr), 256, 0, stream>>>(
                    sg, piece_start, heavy, scratch, scratch_stride);
        }
        if (phases & kStripMerge)
            // two waves per unit: units are 8 rows or more (256 threads: 0.044 vs
            // 0.0415 ms at the 2048^2 bench size, 0.063 vs 0.056 ms clustered)
            strip_merge_kernel<T, NS_><<<dim3((unsigned)sg.nunits, ntr), 128, 0, stream>>>(
                    sg, piece_start, scratch, scratch_stride, fw, fw_stride);
    });
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

template <typename T>
int launch_ragged_strip(const RaggedLaunch &rl, const StripPts<T> &pts, const T *coef,
                        int kerevalmeth, double beta, T *fw_all, T *scratch_all, int phases,
                        hipStream_t stream) {
    // no points anywhere: the caller zeroes the fine grids instead
    if (rl.nv == 0 || pts.M == 0) return JFA_OK;
    if (phases & kStripUnits) {
        const unsigned nblk = (unsigned)((rl.nunits + 255) / 256);
        ragged_units_count_kernel<<<nblk, 256, 0, stream>>>(rl.virt, rl.vunit_off, rl.unit_virt,
                                                            rl.nv, rl.bin_start, rl.piece_start,
                                                            rl.blocksum);
        int *blockoff = rl.blocksum + nblk;  // nblk + 1 entries
        ragged_units_blocks_kernel<<<1, 1024, 0, stream>>>(rl.vunit_off, rl.nv, (int)nblk,
                                                           rl.blocksum, blockoff,
                                                           rl.piece_start, rl.npieces, rl.heavy);
        ragged_units_scan_kernel<<<nblk, 256, 0, stream>>>(rl.vunit_off, rl.nv, blockoff,
                                                           rl.piece_start, rl.piece_unit,
                                                           rl.heavy);
        JFA_HIP_CHECK(hipGetLastError());
    }
    // one workgroup per piece: piece costs differ by problem density, and the
    // hardware dispatcher balances them as waves finish, where a capped grid
    // looping over pieces fixes each workgroup's share up front (64-problem
    // benchmark batch: 4.97 vs 5.31 ms)
    const int64_t ub = rl.nunits + pts.M / rl.pmax + rl.nv;
    const unsigned grid = (unsigned)(ub < (1 << 30) ? ub : (1 << 30));
    const T b = (T)beta;
    JFA_DISPATCH_NS(rl.ns, {
        if (phases & kStripSpread)
            ragged_strip_kernel<T, NS_><<<grid, kStripLanes, 0, stream>>>(
                rl.virt, rl.unit_virt, pts, rl.bin_start, rl.piece_start, rl.npieces,
                rl.piece_unit, coef, b, kerevalmeth, rl.ptrs, rl.nprob, scratch_all);
        if (phases & kStripMerge) {
            const int64_t rb = rl.max_heavy * rl.reduce_chunks;
            ragged_reduce_kernel<T, NS_><<<(unsigned)(rb < 2048 ? rb : 2048), 256, 0, stream>>>(
                rl.virt, rl.unit_virt, rl.piece_start, rl.heavy, rl.reduce_chunks, scratch_all);
        }
        if (phases & kStripMerge)
            ragged_merge_kernel<T, NS_><<<(unsigned)rl.nunits, 64, 0, stream>>>(
                rl.virt, rl.unit_virt, rl.piece_start, scratch_all, fw_all);
    });
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

template <typename T>
int launch_strip3_pack(const T *c, int64_t c_stride, const T *rec3, const T *coef,
                       int kerevalmeth, double beta, int ns, int64_t M, int ntr, T *pack,
                       hipStream_t stream) {
    if (M == 0 || ntr == 0) return JFA_OK;
    const dim3 grid((unsigned)((M + 255) / 256), ntr);
    const T b = (T)beta;
    JFA_DISPATCH_NS(ns, strip3_pack_kernel<T, NS_><<<grid, 256, 0, stream>>>(
                            c, c_stride, rec3, coef, b, kerevalmeth, M, pack));
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

template <typename T>
int launch_strip_pack(const T *c, int64_t c_stride, const T *rec, int64_t M, int ntr, T *pack,
                      hipStream_t stream) {
    if (M == 0 || ntr == 0) return JFA_OK;
    strip_pack_kernel<T><<<dim3((unsigned)((M + 255) / 256), ntr), 256, 0, stream>>>(
        c, c_stride, rec, M, pack);
    JFA_HIP_CHECK(hipGetLastError());
    return JFA_OK;
}

#define JFA_INST(T)                                                                           \
    template int launch_spread_strip<T>(const StripGeom &, const StripPts<T> &, const int *, \
                                        int *, int *, int *, int *, int *, const T *, int,    \
                                        double, int64_t, T *, int64_t, T *, int64_t, int, int, \
                                        hipStream_t);                                         \
    template int launch_strip3_pack<T>(const T *, int64_t, const T *, const T *, int, double, \
                                       int, int64_t, int, T *, hipStream_t);                  \
    template int launch_strip_pack<T>(const T *, int64_t, const T *, int64_t, int, T *,       \
                                      hipStream_t);                                           \
    template int launch_ragged_strip<T>(const RaggedLaunch &, const StripPts<T> &, const T *, \
                                        int, double, T *, T *, int, hipStream_t);
JFA_INST(float)
JFA_INST(double)
#undef JFA_INST

}  // namespace jfa
