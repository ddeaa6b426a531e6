// lu_fused.hpp -- one-workgroup-per-matrix left-looking LU for 64 < n <= 512.
//
// The whole factorisation of a matrix runs in one launch and one workgroup
// of NMAX threads (NMAX = 128, 256 or 512; one matrix row per thread), one
// matrix per workgroup, several workgroups per CU.  Per block column s (width 32):
//   1. gather: the block column is read from the input in its original row
//      order (coalesced, staged through LDS) and placed in MFMA accumulators
//      in the current logical row order;
//   2. left-looking update, for every factored block column t < s:
//        U_ts = L_tt^-1 C[t-rows]        (inv(L_tt) C_t on the matrix cores;
//                                         inv(L_tt) is formed once, right
//                                         after panel t, into a workspace)
//        C[rows below t] -= L21_t * U_ts (MFMA 16x16x4, L21_t staged in LDS in
//                                         16-column stages prefetched one stage
//                                         ahead into registers)
//   3. panel: the accumulators are transposed through LDS to one logical row
//      per thread and factored with partial pivoting, one barrier per column:
//      every wave publishes its best candidate row (arg-max, ties to the
//      lower row -- the reference rule, see pivot_key) and all threads pick
//      the winner;
//   4. the finished block column is written to the output in the logical
//      order of that moment.
// Row interchanges never touch memory.  Rows of L21_t are left where step t
// wrote them and later interchanges are recorded in a per-block map
// (ssig[t][logical row] = storage row), applied when step 2 loads L21_t;
// a last pass moves every L21_t into the final order, so the output is the
// LAPACK layout (A = P L U, 1-based ipiv, info = first zero pivot).
// Output may alias the input (in-place getrf) or be a separate buffer of
// another precision (the FP32 factorisation of an FP64 matrix).
#pragma once

#include "internal.hpp"
#include "mfma.hpp"

#include <type_traits>

namespace rsa
{

constexpr int FNB = 32; // block-column width
constexpr int FHC = 16; // columns per LDS round of the staging/transposition buffer

// Factors one matrix with the whole workgroup (blockDim.x == NMAX, n <= NMAX)
// and returns the first zero pivot (1-based, 0 if none) in every thread.
// With row and column scales (rs, cs non-null) the matrix factored is
// diag(rs) A diag(cs), formed in TI arithmetic as each entry is read.  OPK
// selects the matrix-core operands of the trailing updates: the factor
// precision itself, or FP16 / BF16 operands with FP32 accumulation (TC =
// float; the panel, the block solves and the stored factors stay FP32).
// INVM (inverse mode) 1 or 2: the block solves U = L_tt^-1 C run as
// inv(L_tt) C on the matrix cores.  INVM 1 keeps inv(L_tt) of every block
// column but the last in winv (32x32 column-major blocks, block t at
// winv + 1024 t) and returns the plain factors.  INVM 2 (the internal
// factors of the mixed-precision solvers) returns every 32x32 diagonal block
// inverted instead -- strictly lower part inv(L_tt), upper part with the
// diagonal inv(U_tt) -- and reads inv(L_tt) back from there.  INVM 0: each
// block solve is a
// substitution with the solved values passed between lanes by v_readlane.
template <typename TI, typename TC, int NMAX, int OPK = MFMA_NATIVE, int INVM = 0>
__device__ __forceinline__ int lu_fused_body(int n, const TI* __restrict__ Ai, int ldi,
                                             TC* __restrict__ Ao, int ldo, int* __restrict__ ipiv,
                                             const double* __restrict__ rs = nullptr,
                                             const double* __restrict__ cs = nullptr,
                                             TC* __restrict__ winv = nullptr)
{
    static_assert(OPK == MFMA_NATIVE || std::is_same<TC, float>::value,
                  "half-width operands accumulate in FP32");
    static_assert(INVM >= 0 && INVM <= 2, "inverse mode");
    constexpr bool INVW = INVM != 0;
    using M = mfma16<TC>;
    using acc_t = typename M::acc_t;
    constexpr int NW = NMAX / 64;
    constexpr int SMAX = NMAX / FNB;
    constexpr int SB = FHC + 1; // padded row stride of the row-major view
    constexpr int UB = FNB + 1; // padded row stride of the 32x32 blocks
    // columns of a block column staged per LDS round (half a block: keeps the
    // prefetch registers and LDS -- ~32 KB FP32, ~58 KB FP64 -- small)
    constexpr int KC = FNB / 2;
    constexpr int NKC = FNB / KC;
    // column stride of the gather / L21 staging [col][row]: NMAX + 1, so
    // the 16 columns that one MFMA-layout read takes from a row fall in 16
    // different LDS banks (a stride of NMAX puts them all in one bank)
    constexpr int LS = NMAX + 1;
    // column stride of the L21 staging, whose rows are stored in logical
    // order (the row maps are applied when the block is loaded): a 16-row
    // MFMA-layout read takes 16 consecutive rows of one column, and with
    // 4-byte entries the next lane group's column must start 16 banks on
    // (4 columns on for the K = 16 half-width reads) so that 32 lanes meet
    // 32 different banks
    constexpr int LS2
        = NMAX + (sizeof(TC) == 8 ? 1 : (OPK == MFMA_NATIVE ? 16 : 4));
    constexpr int SBUF = (NMAX * SB > KC * LS2) ? NMAX * SB : KC * LS2;
    constexpr int LTT = FNB * FNB / NMAX; // L_tt elements staged per thread

    __shared__ TC sbuf[SBUF]; // gather / L21 staging [col][row], transposition [row][SB]
    __shared__ TC sU[FNB * UB];
    __shared__ TC sLtt[FNB * UB];
    __shared__ int spi[NMAX]; // logical row -> original row
    __shared__ short ssig[SMAX * NMAX]; // block t: logical row -> storage row
    __shared__ __align__(16) TC scand[2 * NW * FNB];
    __shared__ TC skey[2 * NW];
    __shared__ TC spv[2 * NW];
    __shared__ int slab[2 * NW];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int S = (n + FNB - 1) / FNB;
    // input and factors through buffer resources (branch-free loads, entries
    // past the matrix read 0, stores there are dropped)
    const MatrixBuffer<TI> Ain(Ai, ldi, n);
    const MatrixBuffer<TC> Aout(Ao, ldo, n);

    spi[tid] = tid;
    int zero_at = 0;
    __syncthreads();

    for(int s = 0; s < S; ++s)
    {
        const int c0 = s * FNB;
        const int jb = (n - c0) < FNB ? (n - c0) : FNB;
        acc_t acc[4][2];

        // ---- 1. gather block column s into the accumulators (logical order):
        //         thread tid loads original row tid.  The loads of the second
        //         KC columns are issued as the first KC are consumed, so both
        //         LDS rounds are served by about one memory round trip while
        //         at most KC loads are held in registers.
        {
            uint32_t off = Ain.offset(tid < n, tid, c0);
            const uint32_t step = tid < n ? Ain.ld_bytes : 0u;
            TI a0[KC], a1[KC];
#pragma unroll
            for(int j = 0; j < KC; ++j)
            {
                a0[j] = Ain.load(off, 0); // 0 past row or column n
                off += step;
                opaque(off);
            }
            auto put = [&](int kc, int j, TI v) {
                const int col = c0 + KC * kc + j;
                sbuf[j * LS + tid]
                    = rs ? TC(double(v) * rs[tid < n ? tid : 0] * cs[col < n ? col : 0]) : TC(v);
            };
            auto take = [&](int kc) {
                __syncthreads();
#pragma unroll
                for(int ct = 0; ct < 2; ++ct)
                    if(ct / (2 / NKC) == kc)
#pragma unroll
                        for(int rt = 0; rt < 4; ++rt)
#pragma unroll
                            for(int r = 0; r < 4; ++r)
                            {
                                const int i = 64 * w + 16 * rt + M::row(lane, r);
                                const int j = 16 * ct - KC * kc + (lane & 15);
                                acc[rt][ct][r] = sbuf[j * LS + spi[i]];
                            }
                __syncthreads();
            };
            static_assert(NKC == 2, "two LDS rounds per block column");
#pragma unroll
            for(int j = 0; j < KC; ++j)
            {
                put(0, j, a0[j]);
                a1[j] = Ain.load(off, 0);
                off += step;
                opaque(off);
            }
            take(0);
#pragma unroll
            for(int j = 0; j < KC; ++j)
                put(1, j, a1[j]);
            take(1);
        }

        // ---- 2. left-looking update with every factored block column t < s.
        // Stage (t, kc) = columns [KC*kc, KC*kc+KC) of L21_t (and L_tt when
        // kc == 0).  Its global loads are issued into registers one stage
        // ahead, right after the previous stage's first barrier, so they are
        // in flight during that stage's substitution and MFMAs.
        TC pre[KC];
        TC preL[LTT];
        auto issue = [&](int t, int kc) {
            const int r0 = t * FNB, r1 = r0 + FNB;
            const bool act = tid >= r1 && tid < n;
            // logical row tid of L21_t from its storage row (the block's map)
            const int srow = act ? int(ssig[t * NMAX + tid]) : 0;
            uint32_t off = Aout.offset(act, srow, r0 + KC * kc);
            const uint32_t step = act ? Aout.ld_bytes : 0u;
#pragma unroll
            for(int j = 0; j < KC; ++j)
            {
                pre[j] = Aout.load(off, 0);
                off += step;
                opaque(off);
            }
            if(kc == 0)
#pragma unroll
                for(int e = 0; e < LTT; ++e)
                {
                    const int idx = tid + e * NMAX, i = idx % FNB, k = idx / FNB;
                    if constexpr(INVM == 1)
                        preL[e] = winv[t * FNB * FNB + idx]; // inv(L_tt)(i, k), column-major
                    else if constexpr(INVM == 2)
                    {
                        // strictly lower part of the inverted diagonal block;
                        // unit diagonal, zeros above (lane positions from an
                        // opaque thread id: kept from entry they would spill)
                        const int x2 = fresh_tid() + e * NMAX;
                        const int i2 = x2 % FNB, k2 = x2 / FNB;
                        const TC v = Aout.load(Aout.offset(i2 > k2, r0 + i2, r0 + k2), 0);
                        preL[e] = i2 == k2 ? TC(1) : v;
                    }
                    else
                        preL[e] = Aout.load(Aout.offset(true, r0 + i, r0 + k), 0); // L_tt (final)
                }
        };
        if(s > 0)
            issue(0, 0);
        for(int t = 0; t < s; ++t)
        {
            const int r0 = t * FNB;
            const int r1 = r0 + FNB;
            const int wt = r0 >> 6; // wave holding logical rows [r0, r1)
            const int half = (r0 >> 5) & 1; // which 32-row half of that wave
#pragma unroll
            for(int kc = 0; kc < NKC; ++kc)
            {
#pragma unroll
                for(int j = 0; j < KC; ++j)
                    sbuf[j * LS2 + tid] = pre[j];
                if(kc == 0)
                {
#pragma unroll
                    for(int e = 0; e < LTT; ++e)
                    {
                        const int idx = tid + e * NMAX, i = idx % FNB, k = idx / FNB;
                        sLtt[i * UB + k] = preL[e];
                    }
                    if(w == wt)
                    {
#pragma unroll
                        for(int rt = 0; rt < 4; ++rt)
                            if((rt >> 1) == half)
#pragma unroll
                                for(int ct = 0; ct < 2; ++ct)
#pragma unroll
                                    for(int r = 0; r < 4; ++r)
                                    {
                                        const int q = 16 * (rt & 1) + M::row(lane, r);
                                        sU[q * UB + 16 * ct + (lane & 15)] = acc[rt][ct][r];
                                    }
                    }
                }
                __syncthreads();
                {
                    const int nkc = (kc + 1 == NKC) ? 0 : kc + 1;
                    const int nt = (kc + 1 == NKC) ? t + 1 : t;
                    if(nt < s)
                        issue(nt, nkc);
                }
                if(kc == 0)
                {
                    if constexpr(INVW)
                    {
                        // U_ts = inv(L_tt) C_t on the matrix cores.  Wave w < 2
                        // forms columns [16w, 16w + 16) of U_ts: rows 0-15 from
                        // k < 16 (inv(L_tt) is lower triangular), rows 16-31
                        // from k < 32.  It reads only its own columns of sU
                        // (all of them before it writes), so the result goes
                        // back in place without a further barrier.
                        if(w < 2)
                        {
                            const int c = 16 * w + (lane & 15);
                            const int kq = lane >> 4;
                            const int ri = lane & 15;
                            TC bv[FNB / 4], a0[FNB / 8], a1[FNB / 4];
#pragma unroll
                            for(int ks = 0; ks < FNB / 4; ++ks)
                                bv[ks] = sU[(4 * ks + kq) * UB + c];
#pragma unroll
                            for(int ks = 0; ks < FNB / 8; ++ks)
                                a0[ks] = sLtt[ri * UB + 4 * ks + kq];
#pragma unroll
                            for(int ks = 0; ks < FNB / 4; ++ks)
                                a1[ks] = sLtt[(16 + ri) * UB + 4 * ks + kq];
                            acc_t u0, u1;
#pragma unroll
                            for(int r = 0; r < 4; ++r)
                            {
                                u0[r] = TC(0);
                                u1[r] = TC(0);
                            }
#pragma unroll
                            for(int ks = 0; ks < FNB / 4; ++ks)
                            {
                                if(ks < FNB / 8)
                                    u0 = M::run(a0[ks], bv[ks], u0);
                                u1 = M::run(a1[ks], bv[ks], u1);
                            }
#pragma unroll
                            for(int r = 0; r < 4; ++r)
                            {
                                sU[M::row(lane, r) * UB + c] = u0[r];
                                sU[(16 + M::row(lane, r)) * UB + c] = u1[r];
                            }
                        }
                    }
                    else
                    {
                        // U_ts = L_tt^-1 C_t: lane i of each 32-lane group owns row i
                        // of CPT right-hand columns; x_k reaches the group through
                        // v_readlane (both halves of the wave read, each keeps its own)
                        constexpr int CPT = FNB * FNB / NMAX;
                        const int i = tid & (FNB - 1);
                        const int j0 = (tid / FNB) * CPT;
                        TC x[CPT];
#pragma unroll
                        for(int q = 0; q < CPT; ++q)
                            x[q] = sU[i * UB + j0 + q];
#pragma unroll
                        for(int k = 0; k < FNB - 1; ++k)
                        {
                            const TC lik = sLtt[i * UB + k];
#pragma unroll
                            for(int q = 0; q < CPT; ++q)
                            {
                                const TC x_lo = readlane_t(x[q], k);
                                const TC x_hi = readlane_t(x[q], FNB + k);
                                const TC xk = (lane & FNB) ? x_hi : x_lo;
                                if(i > k)
                                    x[q] = fma_t(-lik, xk, x[q]);
                            }
                        }
#pragma unroll
                        for(int q = 0; q < CPT; ++q)
                            sU[i * UB + j0 + q] = x[q];
                    }
                    __syncthreads();
                    if(w == wt)
                    {
#pragma unroll
                        for(int rt = 0; rt < 4; ++rt)
                            if((rt >> 1) == half)
#pragma unroll
                                for(int ct = 0; ct < 2; ++ct)
#pragma unroll
                                    for(int r = 0; r < 4; ++r)
                                    {
                                        const int q = 16 * (rt & 1) + M::row(lane, r);
                                        acc[rt][ct][r] = sU[q * UB + 16 * ct + (lane & 15)];
                                    }
                    }
                }
                // C[rows >= r1] -= L21_t[:, KC*kc : KC*kc+KC] * U_ts[KC*kc : ..., :]
                if constexpr(OPK != MFMA_NATIVE)
                {
                    // one K = 16 instruction per tile: lane supplies k = kb .. kb+3
                    const int kb = 4 * (lane >> 4);
                    float bv[2][4];
#pragma unroll
                    for(int ct = 0; ct < 2; ++ct)
#pragma unroll
                        for(int e = 0; e < 4; ++e)
                            bv[ct][e] = sU[(KC * kc + kb + e) * UB + 16 * ct + (lane & 15)];
#pragma unroll
                    for(int rt = 0; rt < 4; ++rt)
                    {
                        // wave-uniform; tiles wholly past row n hold zeros that are
                        // never stored
                        if(64 * w + 16 * rt >= r1 && 64 * w + 16 * rt < n)
                        {
                            const int i = 64 * w + 16 * rt + (lane & 15);
                            float av[4];
#pragma unroll
                            for(int e = 0; e < 4; ++e)
                                av[e] = -sbuf[(kb + e) * LS2 + i];
#pragma unroll
                            for(int ct = 0; ct < 2; ++ct)
                                acc[rt][ct] = mfma16x16x16<OPK>(av, bv[ct], acc[rt][ct]);
                        }
                    }
                }
                else
                {
#pragma unroll 2
                    for(int kk = 0; kk < KC; kk += 4)
                    {
                        TC bop[2];
#pragma unroll
                        for(int ct = 0; ct < 2; ++ct)
                            bop[ct] = sU[(KC * kc + kk + (lane >> 4)) * UB + 16 * ct + (lane & 15)];
#pragma unroll
                        for(int rt = 0; rt < 4; ++rt)
                        {
                            // wave-uniform; tiles wholly past row n hold zeros that are
                            // never stored
                            if(64 * w + 16 * rt >= r1 && 64 * w + 16 * rt < n)
                            {
                                const int i = 64 * w + 16 * rt + (lane & 15);
                                const TC aop = -sbuf[(kk + (lane >> 4)) * LS2 + i];
#pragma unroll
                                for(int ct = 0; ct < 2; ++ct)
                                    acc[rt][ct] = M::run(aop, bop[ct], acc[rt][ct]);
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }

        // ---- 3. panel: one logical row per thread
        TC rowv[FNB];
#pragma unroll
        for(int ct = 0; ct < 2; ++ct)
        {
#pragma unroll
            for(int rt = 0; rt < 4; ++rt)
#pragma unroll
                for(int r = 0; r < 4; ++r)
                {
                    const int i = 64 * w + 16 * rt + M::row(lane, r);
                    sbuf[i * SB + (lane & 15)] = acc[rt][ct][r];
                }
            __syncthreads();
#pragma unroll
            for(int j = 0; j < FHC; ++j)
                rowv[FHC * ct + j] = sbuf[tid * SB + j];
            __syncthreads();
        }

        if constexpr(INVM == 2)
        {
            // a partial last diagonal block is inverted as diag(block, I):
            // its two LDS matrices start as 0 (L) and I (U~), the rows of the
            // block overwrite their part after the panel (its barriers order
            // these writes first)
            if(jb < FNB) // uniform
#pragma unroll
                for(int e = 0; e < 2 * FNB * FNB / NMAX; ++e)
                {
                    const int idx = fresh_tid() + e * NMAX; // opaque: not formed at entry
                    const int rem = idx % (FNB * FNB);
                    sbuf[idx] = (idx >= FNB * FNB && rem / FNB == rem % FNB) ? TC(1) : TC(0);
                }
        }
        int label = tid; // logical row held by this thread
#pragma unroll
        for(int k = 0; k < FNB; ++k)
        {
            if(k < jb) // uniform
            {
                const int c = c0 + k;
                const int par = k & 1;
                TC key = (label >= c && label < n) ? pivot_key(rowv[k], label == c) : TC(-1);
                int lab = label;
                wave_argmax(key, lab);
                if(label == lab)
                {
                    // whole 16-byte groups from the one holding column k
                    // (aligned 16-byte stores; entries left of k are unused),
                    // and the candidate's key, row and pivot value
                    constexpr int G = 16 / int(sizeof(TC));
#pragma unroll
                    for(int j = 0; j < FNB; ++j)
                        if(j / G >= k / G)
                            scand[(par * NW + w) * FNB + j] = rowv[j];
                    skey[par * NW + w] = key;
                    slab[par * NW + w] = lab;
                    spv[par * NW + w] = rowv[k];
                }
                __syncthreads();
                // every wavefront's candidate read first, then a branch-free
                // selection (ties to the lower row; bitwise operators, so no
                // short-circuit branches)
                TC kws[NW], pvs[NW];
                int lws[NW];
#pragma unroll
                for(int ww = 0; ww < NW; ++ww)
                {
                    kws[ww] = skey[par * NW + ww];
                    lws[ww] = slab[par * NW + ww];
                    pvs[ww] = spv[par * NW + ww];
                }
                int wb = 0;
                TC bk = kws[0];
                int bl = lws[0];
                TC pv = pvs[0];
#pragma unroll
                for(int ww = 1; ww < NW; ++ww)
                {
                    const bool better = (kws[ww] > bk) | ((kws[ww] == bk) & (lws[ww] < bl));
                    bk = better ? kws[ww] : bk;
                    bl = better ? lws[ww] : bl;
                    pv = better ? pvs[ww] : pv;
                    wb = better ? ww : wb;
                }
                // the winner's row (its index held in a vector register: as a
                // scalar, every read of it waits for the previous one; the
                // row offset stays a visible multiple of FNB, so the reads
                // are 16-byte ds_read_b128 with immediate offsets); the
                // reciprocal of the pivot is formed while those reads are in
                // flight
                uint32_t wbv = uint32_t(wb);
                opaque(wbv);
                const TC* prow = scand + (par * NW) * FNB + wbv * FNB;
                const bool nz = pv != TC(0);
                const TC rp = nz ? TC(1) / pv : TC(1);
                if(!nz && zero_at == 0)
                    zero_at = c + 1;
                if(tid == 0)
                    ipiv[c] = bl + 1;
                // the interchange of logical rows c and bl moves labels only;
                // the row maps take this block column's interchanges at once
                // after the panel
                if(label == bl)
                    label = c;
                else if(label == c)
                    label = bl;
                if(label > c && label < n)
                {
                    TC l = rowv[k];
                    if(nz)
                    {
                        l *= rp;
                        rowv[k] = l;
                    }
                    // 8-column chunks: a chunk's reads are in flight while the
                    // previous chunk's fmas run
                    constexpr int CH = 8;
#pragma unroll
                    for(int c8 = (k + 1) / CH * CH; c8 < FNB; c8 += CH)
                    {
                        // whole 16-byte groups (ds_read_b128; the candidate
                        // row was stored as whole groups from the one
                        // holding column k, and entries left of k are unused)
                        constexpr int G = 16 / int(sizeof(TC));
                        typedef TC g16_t __attribute__((ext_vector_type(G)));
                        TC u[CH];
#pragma unroll
                        for(int g = 0; g < CH / G; ++g)
                        {
                            if(c8 + g * G + G - 1 > k)
                            {
                                const g16_t v = *reinterpret_cast<const g16_t*>(prow + c8 + g * G);
#pragma unroll
                                for(int e = 0; e < G; ++e)
                                    u[g * G + e] = v[e];
                            }
                            else
#pragma unroll
                                for(int e = 0; e < G; ++e)
                                    u[g * G + e] = TC(0);
                        }
#pragma unroll
                        for(int q = 0; q < CH; ++q)
                            if(c8 + q > k)
                                rowv[c8 + q] = fma_t(-l, u[q], rowv[c8 + q]);
                    }
                }
            }
        }

        // ---- 4. write block column s in the logical order of this moment
        //         (with INVM 2 the rows of the diagonal block are written by
        //         the inversion below, so each entry is stored once)
        {
            const bool wr = label < n && (INVM != 2 || label < c0 || label >= c0 + jb);
            uint32_t off = Aout.offset(wr, label, c0);
            const uint32_t step = wr ? Aout.ld_bytes : 0u;
#pragma unroll
            for(int j = 0; j < FNB; ++j)
            {
                Aout.store(rowv[j], off, 0); // columns past n are dropped
                off += step;
                opaque(off);
            }
        }
        // ---- the block column's interchanges reach the row maps in one
        // parallel step: the row at logical position tid when the panel began
        // is now at logical position label, so every map M (spi, and ssig of
        // each earlier block column) becomes M'[label] = M[tid] -- read
        // before the barrier below, written after it.  With INVW the same
        // barrier publishes the rows of L_ss, staged in LDS by columns
        // (sbuf[m][i] = L(i, m), i > m), for the inversion that follows.
        // With INVM 2 the upper triangle goes to a second LDS matrix in its
        // index-reversed lower form U~ (U~(31 - k, 31 - i) = U(i, k), k >= i).
        // uniform; with INVM 1 the last block has no later user
        const bool inv_now = INVM == 2 || (INVM == 1 && s + 1 < S);
        if(inv_now)
        {
            const int i = label - c0;
            if(i >= 0 && i < jb)
            {
#pragma unroll
                for(int m = 0; m < FNB - 1; ++m)
                    if(m < i)
                        sbuf[m * FNB + i] = rowv[m];
                if constexpr(INVM == 2)
#pragma unroll
                    for(int k = 0; k < FNB; ++k)
                        if(k >= i && k < jb)
                            sbuf[FNB * FNB + (FNB - 1 - k) * FNB + (FNB - 1 - i)] = rowv[k];
            }
        }
        {
            const int mspi = spi[tid];
            short msig[SMAX];
#pragma unroll
            for(int t = 0; t < SMAX; ++t)
                if(t < s)
                    msig[t] = ssig[t * NMAX + tid];
            __syncthreads();
            spi[label] = mspi;
#pragma unroll
            for(int t = 0; t < SMAX; ++t)
                if(t < s)
                    ssig[t * NMAX + label] = msig[t];
        }
        if constexpr(INVW)
        {
            // inv(L_ss) for the later block columns' U = inv(L) C: one
            // wavefront solves L X = I by columns -- lane j (< 32) holds
            // column j of X and runs the column-oriented forward substitution,
            // x_i -= L(i, m) x_m for i > m, the L entries read as LDS
            // broadcasts -- and writes the whole 32x32 X (ones on the
            // diagonal, zeros above) column-major to block s of winv.
            if(inv_now && w == s % NW) // one wavefront; the others wait at the next barrier
            {
                // thread id made opaque here: the identity columns would
                // otherwise be formed once at entry and kept live (spilled)
                const int ft = fresh_tid();
                const int j = ft & (FNB - 1);
                // INVM 2: lanes 32..63 invert U~ the same way, dividing by its
                // diagonal (x_m /= d_m; d_m = 1 for L, an exact division)
                const int h = INVM == 2 ? (ft >> 5) & 1 : 0;
                TC x[FNB];
#pragma unroll
                for(int r = 0; r < FNB; ++r)
                    x[r] = r == j ? TC(1) : TC(0);
                uint32_t base = uint32_t(h) * uint32_t(FNB * FNB);
#pragma unroll
                for(int m = 0; m < FNB; ++m)
                {
                    // the column's reads wait for x_m (the empty asm ties
                    // their address to it), or the unrolled solve hoists
                    // all 496 reads to its top
                    asm volatile("" : "+v"(base) : "v"(x[m]));
                    const TC* col = sbuf + base + m * FNB;
                    if constexpr(INVM == 2)
                    {
                        const TC dv = col[m];
                        x[m] = x[m] / (h ? dv : TC(1));
                    }
#pragma unroll
                    for(int r = m + 1; r < FNB; ++r)
                        x[r] = fma_t(-col[r], x[m], x[r]);
                }
                if constexpr(INVM == 1)
                {
                    if(lane < FNB)
                    {
                        // element stores: winv need not be 16-byte aligned
                        TC* dst = winv + s * FNB * FNB + j * FNB;
#pragma unroll
                        for(int r = 0; r < FNB; ++r)
                            dst[r] = x[r];
                    }
                }
                else
                {
                    // results through LDS (every lane has read its matrix)
                    // so the block is stored a column at a time: entry (i, k)
                    // is inv(L)(i, k) below the diagonal, inv(U)(i, k) =
                    // inv(U~)(31 - i, 31 - k) on and above it
                    constexpr int CB = FNB + 1;
                    wave_sync();
#pragma unroll
                    for(int r = 0; r < FNB; ++r)
                    {
                        if(h == 0)
                            sbuf[r * CB + j] = x[r];
                        else
                            sbuf[FNB * CB + (FNB - 1 - r) * CB + (FNB - 1 - j)] = x[r];
                    }
                    wave_sync();
                    const int ft2 = fresh_tid(); // opaque: addresses not formed at entry
                    const int i = ft2 & (FNB - 1);
#pragma unroll
                    for(int t2 = 0; t2 < FNB / 2; ++t2)
                    {
                        const int k = ((ft2 >> 5) & 1) + 2 * t2;
                        const TC v = i > k ? sbuf[i * CB + k] : sbuf[FNB * CB + i * CB + k];
                        Aout.store(v, Aout.offset(i < jb && k < jb, c0 + i, c0 + k), 0);
                    }
                }
            }
        }
        ssig[s * NMAX + tid] = short(tid);
        __syncthreads();
    }

    // ---- move every L21_t into the final row order: row i of block column
    // t takes storage row ssig[t][i].  Each thread loads its source row (32
    // entries), a barrier orders every load of the block column before any
    // store into it, and the next block column's loads are issued before
    // this one's stores, so the pass costs about one memory round trip.
    if(S > 1)
    {
        auto load_block = [&](int t, TC (&x)[FNB]) {
            const int r1 = (t + 1) * FNB;
            const bool act = tid >= r1 && tid < n;
            uint32_t off = Aout.offset(act, act ? int(ssig[t * NMAX + tid]) : 0, t * FNB);
            const uint32_t step = act ? Aout.ld_bytes : 0u;
#pragma unroll
            for(int j = 0; j < FNB; ++j)
            {
                x[j] = Aout.load(off, 0);
                off += step;
                opaque(off);
            }
        };
        TC cur[FNB];
        load_block(0, cur);
        for(int t = 0; t + 1 < S; ++t)
        {
            __syncthreads(); // every thread's loads of block column t are done
            TC nxt[FNB];
            if(t + 2 < S)
                load_block(t + 1, nxt);
            const int r1 = (t + 1) * FNB;
            const bool act = tid >= r1 && tid < n;
            uint32_t off = Aout.offset(act, tid, t * FNB);
            const uint32_t step = act ? Aout.ld_bytes : 0u;
#pragma unroll
            for(int j = 0; j < FNB; ++j)
            {
                Aout.store(cur[j], off, 0);
                off += step;
                opaque(off);
            }
#pragma unroll
            for(int j = 0; j < FNB; ++j)
                cur[j] = nxt[j];
        }
    }
    return zero_at;
}

// Waves per SIMD the register budget is cut for: without a bound the
// compiler spreads into 300-400 registers (1 wave per SIMD).  FP64 fits 2
// waves (170-240 VGPRs); the FP32 factorisation with 256 or 512 threads fits
// 4 (128 VGPRs, no scratch), so four 256-thread workgroups share a CU.
template <typename TI, typename TC, int NMAX, int OPK, int INVM>
__global__ __launch_bounds__(NMAX, (std::is_same<TC, float>::value && OPK == MFMA_NATIVE && NMAX >= 256) ? 4 : 2) void lu_fused_kernel(int n, Batch<TI> AI, int ldi, Batch<TC> AO,
                                                        int ldo, int* __restrict__ ipivA,
                                                        int64_t strideP, int* __restrict__ infoA,
                                                        int bc, const int* __restrict__ skip,
                                                        const double* __restrict__ rs,
                                                        const double* __restrict__ cs,
                                                        int64_t strideS, Batch<TC> WI)
{
    const int64_t b = blockIdx.x;
    if(skip && skip[b] != 0)
        return;
    const int zero_at = lu_fused_body<TI, TC, NMAX, OPK, INVM>(
        n, AI[b], ldi, AO[b], ldo, ipivA + b * strideP, rs ? rs + b * strideS : nullptr,
        cs ? cs + b * strideS : nullptr, INVM == 1 ? WI[b] : nullptr);
    if(threadIdx.x == 0)
        infoA[b] = zero_at;
}

// INVM 1: WI holds lu_fused_inv_elems(n) elements per matrix; INVM 2: WI is
// unused (the inverted diagonal blocks are returned in AO).
template <typename TI, typename TC, int OPK = MFMA_NATIVE, int INVM = 1>
rocblas_status lu_fused_launch(hipStream_t stream, int n, Batch<TI> AI, int ldi, Batch<TC> AO,
                               int ldo, int* ipiv, int64_t strideP, int* info, int bc,
                               Batch<TC> WI, const int* skip, const double* rs = nullptr,
                               const double* cs = nullptr, int64_t strideS = 0)
{
    static_assert(INVM == 1 || INVM == 2, "fixed-size launches keep the inverses");
    if(bc <= 0 || n <= 0)
        return rocblas_status_success;
    if(INVM == 1 && !WI.base && !WI.ptrs)
        return rocblas_status_internal_error; // the inverse blocks need their workspace
    auto go = [&](auto nmax) {
        constexpr int NMAX = decltype(nmax)::value;
        hipLaunchKernelGGL((lu_fused_kernel<TI, TC, NMAX, OPK, INVM>), dim3(bc), dim3(NMAX), 0,
                           stream,
                           n, AI, ldi, AO, ldo, ipiv, strideP, info, bc, skip, rs, cs, strideS,
                           WI);
    };
    if(n <= 128)
        go(std::integral_constant<int, 128>{});
    else if(n <= 256)
        go(std::integral_constant<int, 256>{});
    else
        go(std::integral_constant<int, 512>{});
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

} // namespace rsa
