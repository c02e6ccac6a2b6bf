// vk_kernels.inl -- gfx950 kernels of the periodic pseudospectral step.
//
// Included by vk_inst.hip once per FFT length (compiled with -DVK_L=<len>), so
// every length's kernels build in their own translation unit, in parallel.
//
// Layouts (complex elements, Hp = spectral row pitch, Hp >= N/2+1):
//   state u, N1, N1+N2 acc, stage input : [3][N][N][Hp]   (x, y, z order)
//   A  (after the x pass)               : [f][N][M][Hp]   y-major
//   B  (after the y pass)               : [f][M][M][Hp]
//   C  (after the z pass, aliases A)    : [3][M][M][Hp]
//   D  (after the forward y pass, aliases B) : [3][N][M][Hp]  y-major
// Every strided access touches whole 128-byte rows (T columns x 16 bytes).
#pragma once
#include "vk_fft.hpp"
#include "vk_launch.h"

#include <type_traits>

namespace vk {


// Minimum waves per SIMD asked of the compiler for each kernel class and work
// type (it caps the VGPRs per thread at 512 / waves; 1 = no floor).  Measured
// at 384 (256^3 solver):
//   * FP32 z pass: a floor of 3 lets the compiler fit the pipelined loads in
//     128 VGPRs without scratch, i.e. 4 waves per SIMD (a floor of 4 spills
//     16 B per lane; no floor takes ~200 VGPRs and runs 1.47x slower);
//   * FP64 z VJP pass: a floor of 2 spills 172 B per lane to scratch, yet the
//     reverse step drops 24.3 -> 22.9 ms (unconstrained: 256 VGPRs + 38 AGPRs,
//     one wave per SIMD);
//   * FP64 x/y VJP passes: no floor (256 VGPRs + up to 176 AGPRs); a floor of
//     2 spills ~600 B per lane and is slower (22.8 -> 23.2 ms);
//   * FP32 z VJP pass: no floor (174 VGPRs, two waves); a floor of 3 (44 B
//     spilled) measured the same;
//   * forward/stage passes: a floor of 2 keeps the stage-3/4 epilogues at two
//     waves (FP64: no AGPR overflow; FP32 stage 3 of the fused pass: 256
//     VGPRs and 17 B spilled instead of 256 VGPRs + 19 AGPRs at one wave).
enum KernelClass { K_INV, K_FWD, K_FVJP, K_ZNL, K_ZVJP };
template <int K, typename TW>
struct Waves {
    static constexpr int value = sizeof(TW) == 8 ? (K == K_FWD ? 2 : K == K_ZVJP ? 2 : 1)
                                                 : (K == K_ZNL ? 3 : K == K_FWD ? 2 : 1);
};

__device__ __forceinline__ long long col_addr(int col, int inner, long long os) {
    return (long long)(col / inner) * os + (col % inner);
}

// Output transforms of the strided inverse passes are stored by every column
// of the tile, padding columns ((col % hp) >= hvalid) included: their values
// land in row padding, which no pass reads.  This needs whole tiles
// (ncols % T == 0), which holds when T divides the row alignment (every
// strided geometry spans whole padded rows); tiles widened for short lines
// keep predicated stores, and the launchers check.  A store under a
// per-column predicate ends in a branch join where the compiler merges its
// outstanding-memory counts conservatively, and the next field's prefetched
// loads are then waited for together with every store of this field (loads
// and stores share one in-order counter).
template <typename TW, int T>
constexpr bool pad_stores() {
    return (VK_HP_ALIGN_BYTES / (2 * (int)sizeof(TW))) % T == 0;
}

// ===========================================================================
// Inverse strided pass: lines of LIN spectral inputs zero-padded to L, inverse
// DFT, all L physical outputs stored.  Curl terms are formed on the way in:
// i*k multiplications commute with transforms along the other axes, so the x
// pass forms i kx (.) and the y pass forms i ky (.) and i kz (.).  Each output
// field re-reads its input fields (repeat reads of a tile hit the L2/MALL).
// ===========================================================================
// single-field modes use the low-pass-count plan; modes that combine fields
// use the smaller-E plan to stay within the register budget
template <int L, int MODE, typename TW>
struct InvPlan {
    using type = typename PlanSmall<L>::type;
};
template <int L, typename TW>
struct InvPlan<L, INV_PLAIN, TW> {
    using type = PlanS<L>;
};
// INV_U3 (the CFL pass's x and y inverses, and the VJPs' state passes): the
// low-pass-count plan in FP32; in FP64 the smaller-E plan, which leaves room
// to load each next field during a transform (the CFL pass at 256: 152 ->
// 112 VGPRs, 3 -> 4 waves per SIMD; FP64 step 112.4 -> 112.9 steps/s with the
// CFL pass overlapping stage 1, same boxes; the VJPs' 384-point instance
// 248 -> 150 VGPRs, reverse step unchanged)
template <int L, typename TW>
struct InvPlan<L, INV_U3, TW> {
    using type = std::conditional_t<sizeof(TW) == 8, typename PlanSmall<L>::type, PlanS<L>>;
};
// the FP32 y inverse runs the 24-element plan too (4-wave blocks of 16
// columns; 0.362 -> 0.339 ms at 384), and prefetches like the 12-element one
template <int L, typename TW>
struct InvPlan<L, INV_UOMEGA, TW> {
    using type = std::conditional_t<sizeof(TW) == 4, PlanS<L>, typename PlanSmall<L>::type>;
};
// The stage passes' x forward plan (k_fwd FWD_STAGE and the fused pass).  FP64
// uses the 12-element plan (2 waves per SIMD at 256 VGPRs); FP32 uses the
// low-pass-count plan (24 elements, radices 8 8 6 at 384): its 16-column
// tiles then make 4-wave blocks instead of 8-wave ones, so two blocks share a
// CU, and one LDS exchange fewer per transform.
template <int L, typename TW>
struct StagePlan {
    using type = std::conditional_t<sizeof(TW) == 4, PlanS<L>, PlanE12<L>>;
};
// The padded x inverse runs the stage plan in reversed radix order, the
// order of the fused stage pass's inverse (k_fwd_stage_inv), so the fused and
// unfused paths round identically.
template <typename PLAN>
struct ReversedPlan {
    using radices = typename Reverse<typename PLAN::radices>::type;
    static constexpr int E = PLAN::E;
};
template <int L, typename TW, bool PADDED = (L % 3 == 0)>
struct XPassPlan {
    using type = ReversedPlan<typename StagePlan<L, TW>::type>;
};
template <int L, typename TW>
struct XPassPlan<L, TW, false> {
    using type = typename PlanSmall<L>::type;
};
template <int L, typename TW>
struct InvPlan<L, INV_U3DX, TW> {
    using type = typename XPassPlan<L, TW>::type;
};
// bytes of each position row of a strided tile (T = row bytes / element bytes
// columns): 128 in every strided pass.  Wider rows need tiles of 98 KB (one
// block per CU) and measured slower.
constexpr int kRowBytes = 128;
template <int L, int MODE, typename TW>
using InvGeom = Geom<L, InvPlan<L, MODE, TW>::type::E, TW, kRowBytes>;

template <int L, int LIN, int MODE, typename TI, typename TW>
__global__ __launch_bounds__((InvGeom<L, MODE, TW>::NT)) __attribute__((amdgpu_waves_per_eu(Waves<K_INV, TW>::value))) void k_inv(const cpx<TI> *__restrict__ in,
                                                              cpx<TW> *__restrict__ out,
                                                              const cpx<TW> *__restrict__ tw, Strided g, int n,
                                                              double ks, TW scale) {
    using P = typename InvPlan<L, MODE, TW>::type;
    using PL = typename P::radices;
    constexpr int E = P::E;
    constexpr int T = InvGeom<L, MODE, TW>::T, NT = InvGeom<L, MODE, TW>::NT;
    // prefetch the next input field while transforming (multi-field modes with
    // the 12-element plan, and the FP32 x pass: a 24-element FP64 prefetch
    // buffer would cost occupancy)
    constexpr bool PF = MODE != INV_PLAIN && (E <= 12 || (sizeof(TW) == 4 && (MODE == INV_U3DX || MODE == INV_UOMEGA)));
    constexpr int TWN = PF ? tw_entries<L, PL>() : 1;
    __shared__ cpx<TW> lds[L * T];
    __shared__ cpx<TW> ltw[TWN];
    const int tid = threadIdx.x;
    const int col0 = (g.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x) * T;
    const cpx<TW> zero = {TW(0), TW(0)};
    cpx<TW> v[E];
    TwStage<TWN, NT, TW> tws;
    if constexpr (PF) twiddles_load(tws, tw, tid);
    // the LDS twiddle writes follow the first field loads (see twiddles_load)
    auto commit = [&] {
        if constexpr (PF) twiddles_store(ltw, tws, tid);
    };

    // this thread's column (NT is a multiple of T) and its base offsets
    const int col = col0 + tid % T;
    const bool colok = col < g.ncols && (col % g.hp) < g.hvalid;
    const long long cin = colok ? col_addr(col, g.in_inner, g.in_os) : 0;
    constexpr bool PADST = pad_stores<TW, T>();
    const long long cout = (PADST || colok) ? col_addr(col, g.out_inner, g.out_os) : 0;
    const TW kz_col = TW(double(col % g.hp) * ks);  // kz of the column (y pass)
    // element e -> unpadded input index (-1 in the zero gap)
    auto uidx = [&](int e) -> int {
        const int p = ci_pos_in<L, PL::first, T, NT>(tid, e);
        return LIN == L ? p : unpad_index(p, LIN, L);
    };
    // A, B: input fields being combined; Q: the next input field, loaded
    // while the current one is transformed (the transforms take their
    // twiddles from LDS, so no vector memory wait inside them holds for Q)
    cpx<TW> A[E], B[E], Q[E];
    // fill a register slot with input field f (opaque pointer: the loads stay
    // where they are written, ahead of the transform they overlap).  Every
    // load is unconditional, from a valid address (column 0 for columns
    // outside the geometry, index 0 for positions in the zero gap), and zeros
    // are selected afterwards: a load under a branch is waited for before the
    // branch joins, which would serialise the loads.  Elements whose whole
    // position range lies in the gap are zero at compile time and never loaded.
    auto fill = [&](cpx<TW>(&S)[E], int f) {
        const gptr<cpx<TI>> src = opaque(in) + f * g.in_fs + cin;
        static_for<E>([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            constexpr int gc = gap_class(ci_in_base<L, PL::first, T, NT>(e), NT / T, LIN, L);
            if constexpr (gc == 1) {
                S[e] = zero;
            } else if constexpr (gc == 0) {
                S[e] = cvt<TW>(ldg(src, (long long)uidx(e) * g.in_ls));
            } else {
                const int i = uidx(e);
                const cpx<TW> x = cvt<TW>(ldg(src, (long long)(i < 0 ? 0 : i) * g.in_ls));
                S[e] = i < 0 ? zero : x;
            }
        });
    };
    auto take = [&](cpx<TW>(&S)[E]) {
#pragma unroll
        for (int e = 0; e < E; ++e) S[e] = Q[e];
    };
    auto kline = [&](int e) -> TW {
        const int i = uidx(e);
        return TW(double(wavenum(i < 0 ? 0 : i, LIN == L ? n : LIN)) * ks);
    };
    auto emit = [&](int fo, auto value) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = value(e);
        if constexpr (PF) fft_reg_lt<L, +1, T, NT, T, 1, PL>(v, lds, ltw, tid);
        else fft_reg<L, +1, T, NT, T, 1, PL>(v, lds, tw, tid);
        if (PADST || colok) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int p = ci_pos_out<L, PL::last, T, NT>(tid, e);
                out[fo * g.out_fs + cout + (long long)p * g.out_ls] = cscale(v[e], scale);
            }
        }
        __syncthreads();
    };
    auto ik = [](TW k, cpx<TW> a) -> cpx<TW> { return {-k * a.im, k * a.re}; };
    auto copyA = [&](int e) { return A[e]; };
    auto copyB = [&](int e) { return B[e]; };
    // y pass curl terms; inputs 0 ux, 1 uy, 2 uz, 3 ax = i kx uz, 4 bx = i kx uy
    auto wx = [&](int e) {  // i (ky uz - kz uy), A = uy, B = uz
        const TW ky = kline(e);
        return ik(TW(1), cpx<TW>{ky * B[e].re - kz_col * A[e].re, ky * B[e].im - kz_col * A[e].im});
    };
    auto wy = [&](int e) {  // i kz ux - ax, A = ux, B = ax
        return cpx<TW>{-kz_col * A[e].im - B[e].re, kz_col * A[e].re - B[e].im};
    };
    auto wz = [&](int e) {  // bx - i ky ux, A = ux, B = bx
        const TW ky = kline(e);
        return cpx<TW>{B[e].re + ky * A[e].im, B[e].im - ky * A[e].re};
    };

    if constexpr (MODE == INV_PLAIN) {
        fill(A, 0);
        emit(0, copyA);
    } else if constexpr (MODE == INV_U3 && !PF) {
#pragma unroll 1
        for (int f = 0; f < 3; ++f) {
            fill(A, f);
            emit(f, copyA);
        }
    } else if constexpr (MODE == INV_U3) {
        fill(A, 0);
        fill(Q, 1);
        commit();
        emit(0, copyA);
        take(A);
        fill(Q, 2);
        emit(1, copyA);
        take(A);
        emit(2, copyA);
    } else if constexpr (MODE == INV_U3DX) {
        static_assert(PF, "the x pass prefetches");
        // outputs: 0 ux, 1 uy, 2 uz, 3 i kx uz, 4 i kx uy
        fill(A, 0);
        fill(Q, 1);
        commit();
        emit(0, copyA);
        take(A);
        fill(Q, 2);
        emit(1, copyA);
        emit(4, [&](int e) { return ik(kline(e), A[e]); });
        take(A);
        emit(2, copyA);
        emit(3, [&](int e) { return ik(kline(e), A[e]); });
    } else {
        // UOMEGA outputs 0..2 u, 3 wx, 4 wy, 5 wz; OMEGA outputs 0 wx, 1 wy, 2 wz.
        // wx is emitted before the plain uz copy: uy is dead once wx's input is
        // formed, so no transform runs with two input fields live besides Q.
        constexpr int OW = (MODE == INV_UOMEGA) ? 3 : 0;
        fill(A, 0);
        fill(Q, 3);
        commit();
        if constexpr (MODE == INV_UOMEGA) emit(0, copyA);
        take(B);
        fill(Q, 4);
        emit(OW + 1, wy);
        take(B);
        fill(Q, 1);
        emit(OW + 2, wz);
        take(A);
        fill(Q, 2);
        if constexpr (MODE == INV_UOMEGA) emit(1, copyA);
        take(B);
        emit(OW + 0, wx);
        if constexpr (MODE == INV_UOMEGA) emit(2, copyB);
    }
}

// ===========================================================================
// Forward strided pass: lines of L physical values, forward DFT, the LOUT kept
// modes (3/2-rule truncation map) scaled and stored.  FWD_STAGE additionally
// projects (Leray) and runs one ETD-RK4 stage epilogue per mode; FWD_VJPY
// (y pass) and FWD_VJP (x pass) form t = GU + curl GW between them, and
// FWD_VJP runs one reverse-stage epilogue per mode.  The stage is
// a template parameter, so each epilogue's loads are known at compile time and
// are issued together per chunk of CH kept modes (the state arrays may alias
// as far as the compiler knows, so loads of the next chunk wait behind the
// stores of this one).  Loads are unconditional (columns outside the geometry
// read column 0); only the stores are predicated.
// ===========================================================================
// ETD-RK4 stage epilogue (vonkarman-periodic/src/solver.rs:264-378) for CH
// modes; nn = Leray-projected N-hat of the modes.  Table rows (etd_index):
// e, e_half, a21 (= a31), a41, b1, b23, b4.
// XO: also return the stage output (the next stage input, or the new state)
// in xo; it is stored to sa.out only when sa.out is set (uniform branch).
// k2[i]: |k|^2 of mode i in FP64 (kx^2 + ky^2 + kz^2, as SpectralOps::k_mag_sq),
// used by the RK4 stages' viscous term.
// The stage passes stream (non-temporal loads and stores) their input
// fields, the state arrays of the epilogues and the fused pass's x-inverse
// outputs: each element is touched once per pass, and streaming keeps the
// ETD table's lines cached (fused pass 0.70 -> 0.66 ms, stage 4 0.35 -> 0.33
// ms at 384 FP64).
template <typename R>
__device__ __forceinline__ cpx<R> ldnt(const cpx<R> *p) {
    using V = R __attribute__((ext_vector_type(2)));
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
    return {v.x, v.y};
}
template <typename R>
__device__ __forceinline__ void stnt(cpx<R> *p, cpx<R> v) {
    using V = R __attribute__((ext_vector_type(2)));
    const V x = {v.re, v.im};
    __builtin_nontemporal_store(x, reinterpret_cast<V *>(p));
}
// x, held at this point in the program: an empty asm that takes and returns
// the value, so nothing it depends on is moved below it (into a branch), and
// nothing computed from it is merged with values computed before it
template <typename R>
__device__ __forceinline__ R pinned(R x) {
    asm volatile("" : "+v"(x));
    return x;
}
// The operands a chunk of CH modes reads from memory: up to four ETD table
// entries and up to three state arrays (three fields each) per mode; which
// ones depends on the stage.  stage_load issues the loads, stage_apply forms
// and stores the results; stage_chunk runs the two in turn.  (Issuing each
// next chunk's loads before this chunk's stores spilled the fused pass at two
// waves per SIMD in every stage.)
template <int CH, typename TS>
struct StageOps {
    double t[4][CH];
    cpx<TS> s[3][CH][3];
};
template <int STG, int CH, typename TS>
__device__ __forceinline__ void stage_load(const StageArgs &sa, const long long (&idx)[CH], const int (&q)[CH],
                                           StageOps<CH, TS> &o) {
    const cpx<TS> *u = reinterpret_cast<const cpx<TS> *>(sa.u);
    const cpx<TS> *n1 = reinterpret_cast<const cpx<TS> *>(sa.n1);
    const cpx<TS> *sacc = reinterpret_cast<const cpx<TS> *>(sa.sacc);
    const cpx<TS> *xin = reinterpret_cast<const cpx<TS> *>(sa.xin);
    const double *tab = sa.tab;
    const int ts = sa.tab_stride;
    const long long fs = sa.fs;
#pragma unroll
    for (int i = 0; i < CH; ++i) {
        if constexpr (STG >= RK_1 && STG <= RK_4) {
            // s0 the stage input, s1 u, s2 N1 (RK_2) or the accumulator (RK_3, RK_4)
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                o.s[0][i][f] = ldnt(xin + f * fs + idx[i]);
                if constexpr (STG != RK_1) o.s[1][i][f] = ldnt(u + f * fs + idx[i]);
                if constexpr (STG == RK_2) o.s[2][i][f] = ldnt(n1 + f * fs + idx[i]);
                if constexpr (STG == RK_3 || STG == RK_4) o.s[2][i][f] = ldnt(sacc + f * fs + idx[i]);
            }
        } else if constexpr (STG == ST_1 || STG == ST_2) {
            o.t[0][i] = tab[etd_index(ETD_EH, q[i], ts)];
            o.t[1][i] = tab[etd_index(ETD_A21, q[i], ts)];
#pragma unroll
            for (int f = 0; f < 3; ++f) o.s[0][i][f] = ldnt(u + f * fs + idx[i]);
        } else if constexpr (STG == ST_3) {
            o.t[0][i] = tab[etd_index(ETD_E, q[i], ts)];
            o.t[1][i] = tab[etd_index(ETD_A41, q[i], ts)];
            o.t[2][i] = tab[etd_index(ETD_B1, q[i], ts)];
            o.t[3][i] = tab[etd_index(ETD_B23, q[i], ts)];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                o.s[0][i][f] = ldnt(u + f * fs + idx[i]);
                o.s[1][i][f] = ldnt(sacc + f * fs + idx[i]);
                o.s[2][i][f] = ldnt(n1 + f * fs + idx[i]);
            }
        } else if constexpr (STG == ST_4) {
            o.t[0][i] = tab[etd_index(ETD_B4, q[i], ts)];
#pragma unroll
            for (int f = 0; f < 3; ++f) o.s[0][i][f] = ldnt(sacc + f * fs + idx[i]);
        }
    }
}
template <int STG, int CH, typename TS, bool XO = false>
__device__ __forceinline__ void stage_apply(const StageArgs &sa, bool st, const long long (&idx)[CH],
                                            const cpx<TS> (&nn)[CH][3], const StageOps<CH, TS> &o,
                                            cpx<TS> (*xo)[3] = nullptr, const double *k2 = nullptr) {
    cpx<TS> *out = reinterpret_cast<cpx<TS> *>(sa.out);
    cpx<TS> *n1 = reinterpret_cast<cpx<TS> *>(sa.n1);
    cpx<TS> *sacc = reinterpret_cast<cpx<TS> *>(sa.sacc);
    const long long fs = sa.fs;
    if constexpr (STG == ST_RAW) {
        if (st) {
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int f = 0; f < 3; ++f) stnt(out + f * fs + idx[i], nn[i][f]);
        }
    } else if constexpr (STG >= RK_1 && STG <= RK_4) {
        // classical RK4 (vonkarman-periodic/src/rk4.rs:30-143): r = N(state) -
        // nu k^2 state; RK_1: N1 <- r1, x <- u + dt/2 r1; RK_2: acc <- N1 +
        // 2 r2, x <- u + dt/2 r2; RK_3: acc <- acc + 2 r3, x <- u + dt r3;
        // RK_4: u <- u + dt/6 (acc + r4)
        const double dt = sa.rk_dt, nu = sa.nu;
        const auto &stv = o.s[0], &uu = o.s[1], &ac = o.s[2];
        if (st) {
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const double visc = -nu * k2[i];
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const double rr = double(nn[i][f].re) + visc * double(stv[i][f].re);
                    const double ri = double(nn[i][f].im) + visc * double(stv[i][f].im);
                    const long long j = f * fs + idx[i];
                    if constexpr (STG == RK_1) {
                        stnt(n1 + j, cpx<TS>{TS(rr), TS(ri)});
                        stnt(out + j, cpx<TS>{TS(stv[i][f].re + 0.5 * dt * rr), TS(stv[i][f].im + 0.5 * dt * ri)});
                    } else if constexpr (STG == RK_2) {
                        stnt(sacc + j, cpx<TS>{TS(ac[i][f].re + 2.0 * rr), TS(ac[i][f].im + 2.0 * ri)});
                        stnt(out + j, cpx<TS>{TS(uu[i][f].re + 0.5 * dt * rr), TS(uu[i][f].im + 0.5 * dt * ri)});
                    } else if constexpr (STG == RK_3) {
                        stnt(sacc + j, cpx<TS>{TS(ac[i][f].re + 2.0 * rr), TS(ac[i][f].im + 2.0 * ri)});
                        stnt(out + j, cpx<TS>{TS(uu[i][f].re + dt * rr), TS(uu[i][f].im + dt * ri)});
                    } else {
                        const double sixth_dt = dt / 6.0;
                        stnt(out + j, cpx<TS>{TS(uu[i][f].re + sixth_dt * (ac[i][f].re + rr)),
                                  TS(uu[i][f].im + sixth_dt * (ac[i][f].im + ri))});
                    }
                }
            }
        }
    } else if constexpr (STG == ST_1 || STG == ST_2) {
        const double dt = sa.sc->dt;
        const auto &eh = o.t[0], &a = o.t[1];
        const auto &uu = o.s[0];
        if (st) {
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<TS> x = nn[i][f];
                    stnt((STG == ST_1 ? n1 : sacc) + f * fs + idx[i], x);
                }
        }
#pragma unroll
        for (int i = 0; i < CH; ++i)
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cpx<TS> x = nn[i][f];
                const cpx<TS> y = {TS(eh[i] * uu[i][f].re + dt * a[i] * x.re), TS(eh[i] * uu[i][f].im + dt * a[i] * x.im)};
                if constexpr (XO) xo[i][f] = y;
                if (st && (!XO || out)) stnt(out + f * fs + idx[i], y);
            }
    } else if constexpr (STG == ST_3) {
        // stage c, and the part of the final combine that does not involve N4:
        // sacc <- e u + dt (b1 N1 + b23 (N2 + N3)), so stage 4 reads one array
        const double dt = sa.sc->dt;
        double b1[CH], b23[CH];
        cpx<TS> s2[CH][3];
        const auto &e = o.t[0], &a = o.t[1];
        const auto &uu = o.s[0], &m1 = o.s[2];
        // b1, b23 and N2 enter only the accumulator, whose store is
        // predicated; held here, their loads go out with the chunk's others
        // instead of after those have returned (one memory round trip per
        // chunk, not two)
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            b1[i] = pinned(o.t[2][i]);
            b23[i] = pinned(o.t[3][i]);
#pragma unroll
            for (int f = 0; f < 3; ++f) s2[i][f] = {pinned(o.s[1][i][f].re), pinned(o.s[1][i][f].im)};
        }
        if (st) {
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<TS> x = nn[i][f];
                    const cpx<TS> s3 = {s2[i][f].re + x.re, s2[i][f].im + x.im};  // N2 + N3
                    const double rr = b1[i] * m1[i][f].re + b23[i] * s3.re;
                    const double ri = b1[i] * m1[i][f].im + b23[i] * s3.im;
                    stnt(sacc + f * fs + idx[i], cpx<TS>{TS(e[i] * uu[i][f].re + dt * rr), TS(e[i] * uu[i][f].im + dt * ri)});
                }
        }
#pragma unroll
        for (int i = 0; i < CH; ++i)
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cpx<TS> x = nn[i][f];
                const double dr = 2.0 * x.re - m1[i][f].re, di = 2.0 * x.im - m1[i][f].im;
                const cpx<TS> y = {TS(e[i] * uu[i][f].re + dt * a[i] * dr), TS(e[i] * uu[i][f].im + dt * a[i] * di)};
                if constexpr (XO) xo[i][f] = y;
                if (st && (!XO || out)) stnt(out + f * fs + idx[i], y);
            }
    } else {  // ST_4: final combine into u, from the stage-3 partial in sacc
        static_assert(STG == ST_4, "unknown stage");
        const double dt = sa.sc->dt;
        const auto &b4 = o.t[0];
        const auto &p = o.s[0];
#pragma unroll
        for (int i = 0; i < CH; ++i)
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cpx<TS> x = nn[i][f];
                const cpx<TS> y = {TS(double(p[i][f].re) + dt * (b4[i] * x.re)),
                                   TS(double(p[i][f].im) + dt * (b4[i] * x.im))};
                if constexpr (XO) xo[i][f] = y;
                if (st && (!XO || out)) stnt(out + f * fs + idx[i], y);
            }
    }
}
// one chunk's loads, then its results
template <int STG, int CH, typename TS, bool XO = false>
__device__ __forceinline__ void stage_chunk(const StageArgs &sa, bool st, const long long (&idx)[CH],
                                            const int (&q)[CH], const cpx<TS> (&nn)[CH][3],
                                            cpx<TS> (*xo)[3] = nullptr, const double *k2 = nullptr) {
    StageOps<CH, TS> o;
    stage_load<STG, CH, TS>(sa, idx, q, o);
    stage_apply<STG, CH, TS, XO>(sa, st, idx, nn, o, xo, k2);
}

// Leray projection of one mode (vonkarman-core/src/spectral_ops.rs:114-151)
__device__ __forceinline__ void leray_mode(double kx, double ky, double kz, cpx<double> x[3]) {
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 < 1e-30) return;
    const double inv = 1.0 / k2;
    const double dr = kx * x[0].re + ky * x[1].re + kz * x[2].re;
    const double di = kx * x[0].im + ky * x[1].im + kz * x[2].im;
    x[0] = {x[0].re - kx * dr * inv, x[0].im - kx * di * inv};
    x[1] = {x[1].re - ky * dr * inv, x[1].im - ky * di * inv};
    x[2] = {x[2].re - kz * dr * inv, x[2].im - kz * di * inv};
}

// Reverse sweep of etd_rk4_step (vonkarman-periodic/src/solver.rs:221-379):
// with t = N'(stage state)^T (cotangent of that stage's N), the stage
// cotangents propagate as
//   AS_4 (stage c):  gu = e (lam + t);  g1p = dt (b1 lam - a41 t);  cot <- dt (b23 lam + 2 a41 t)
//   AS_3 (stage b):  gu += e_half t;    cot <- dt (b23 lam + a31 t)
//   AS_2 (stage a):  gu += e_half t;    cot <- g1p + dt a21 t
//   AS_1 (stage u):  out = gu + t
// where cot <- x stores inv_scale * P(x), the input the next VJP's inverse
// passes expect.  All coefficient operators are real diagonal, hence self-adjoint.
template <int STG, int CH>
__device__ __forceinline__ void adj_chunk(const StageArgs &sa, bool st, const long long (&idx)[CH], const int (&q)[CH],
                                          const double (&kv)[CH][3], const cpx<double> (&t)[CH][3]) {
    const long long fs = sa.fs;
    const cpx<double> *lam = reinterpret_cast<const cpx<double> *>(sa.lam);
    cpx<double> *gu = reinterpret_cast<cpx<double> *>(sa.gu);
    cpx<double> *g1p = reinterpret_cast<cpx<double> *>(sa.g1p);
    cpx<double> *cot = reinterpret_cast<cpx<double> *>(sa.cot);
    cpx<double> *out = reinterpret_cast<cpx<double> *>(sa.out);
    const double *tab = sa.tab;
    const int ts = sa.tab_stride;
    if constexpr (STG == AS_RAW) {
        if (st) {
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int f = 0; f < 3; ++f) out[f * fs + idx[i]] = t[i][f];
        }
        return;
    } else if constexpr (STG == AS_1) {
        cpx<double> g[CH][3];
#pragma unroll
        for (int i = 0; i < CH; ++i)
#pragma unroll
            for (int f = 0; f < 3; ++f) g[i][f] = gu[f * fs + idx[i]];
        if (st) {
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int f = 0; f < 3; ++f)
                    out[f * fs + idx[i]] = {g[i][f].re + t[i][f].re, g[i][f].im + t[i][f].im};
        }
        return;
    } else {
        const double dt = sa.sc->dt;
        cpx<double> c[CH][3];
        if constexpr (STG == AS_4) {
            double e[CH], a41[CH], b1[CH], b23[CH];
            cpx<double> l[CH][3];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                e[i] = tab[etd_index(ETD_E, q[i], ts)];
                a41[i] = tab[etd_index(ETD_A41, q[i], ts)];
                b1[i] = tab[etd_index(ETD_B1, q[i], ts)];
                b23[i] = tab[etd_index(ETD_B23, q[i], ts)];
#pragma unroll
                for (int f = 0; f < 3; ++f) l[i][f] = lam[f * fs + idx[i]];
            }
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<double> tt = t[i][f], ll = l[i][f];
                    if (st) {
                        gu[f * fs + idx[i]] = {e[i] * (ll.re + tt.re), e[i] * (ll.im + tt.im)};
                        g1p[f * fs + idx[i]] = {dt * (b1[i] * ll.re - a41[i] * tt.re),
                                                dt * (b1[i] * ll.im - a41[i] * tt.im)};
                    }
                    c[i][f] = {dt * (b23[i] * ll.re + 2.0 * a41[i] * tt.re), dt * (b23[i] * ll.im + 2.0 * a41[i] * tt.im)};
                }
        } else if constexpr (STG == AS_3) {
            double eh[CH], a31[CH], b23[CH];
            cpx<double> l[CH][3], g[CH][3];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                eh[i] = tab[etd_index(ETD_EH, q[i], ts)];
                a31[i] = tab[etd_index(ETD_A21, q[i], ts)];
                b23[i] = tab[etd_index(ETD_B23, q[i], ts)];
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    l[i][f] = lam[f * fs + idx[i]];
                    g[i][f] = gu[f * fs + idx[i]];
                }
            }
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<double> tt = t[i][f];
                    if (st) gu[f * fs + idx[i]] = {g[i][f].re + eh[i] * tt.re, g[i][f].im + eh[i] * tt.im};
                    c[i][f] = {dt * (b23[i] * l[i][f].re + a31[i] * tt.re), dt * (b23[i] * l[i][f].im + a31[i] * tt.im)};
                }
        } else {
            static_assert(STG == AS_2, "unknown reverse stage");
            double eh[CH], a21[CH];
            cpx<double> g[CH][3], p[CH][3];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                eh[i] = tab[etd_index(ETD_EH, q[i], ts)];
                a21[i] = tab[etd_index(ETD_A21, q[i], ts)];
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    g[i][f] = gu[f * fs + idx[i]];
                    p[i][f] = g1p[f * fs + idx[i]];
                }
            }
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<double> tt = t[i][f];
                    if (st) gu[f * fs + idx[i]] = {g[i][f].re + eh[i] * tt.re, g[i][f].im + eh[i] * tt.im};
                    c[i][f] = {p[i][f].re + dt * a21[i] * tt.re, p[i][f].im + dt * a21[i] * tt.im};
                }
        }
        if (st) {
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                leray_mode(kv[i][0], kv[i][1], kv[i][2], c[i]);
#pragma unroll
                for (int f = 0; f < 3; ++f)
                    cot[f * fs + idx[i]] = {c[i][f].re * sa.inv_scale, c[i][f].im * sa.inv_scale};
            }
        }
    }
}

template <int L, int MODE, typename TW>
struct FwdPlan {
    using type = PlanS<L>;
};
template <int L, typename TW>
struct FwdPlan<L, FWD_STAGE, TW> {
    using type = typename StagePlan<L, TW>::type;
};
template <int L, typename TW>
struct FwdPlan<L, FWD_VJP, TW> {
    using type = PlanE12<L>;
};
// the VJP y pass holds one field of kept outputs beside the transform
template <int L, typename TW>
struct FwdPlan<L, FWD_VJPY, TW> {
    using type = PlanE12<L>;
};
template <int L, int MODE, typename TW>
using FwdGeom = Geom<L, FwdPlan<L, MODE, TW>::type::E, TW, kRowBytes>;

template <int L, int LOUT, int MODE, int NF, int STG, typename TW>
__global__ __launch_bounds__((FwdGeom<L, MODE, TW>::NT)) __attribute__((amdgpu_waves_per_eu(Waves<MODE == FWD_VJP ? K_FVJP : K_FWD, TW>::value))) void k_fwd(const cpx<TW> *__restrict__ in,
                                                                    cpx<TW> *__restrict__ out,
                                                                    const cpx<TW> *__restrict__ tw, Strided g,
                                                                    int n, double ks, TW scale, StageArgs sa) {
    using P = typename FwdPlan<L, MODE, TW>::type;
    using PL = typename P::radices;
    constexpr int E = P::E;
    constexpr int T = FwdGeom<L, MODE, TW>::T, NT = FwdGeom<L, MODE, TW>::NT;
    constexpr int RP = PL::last;
    // the stage pass loads its next field while transforming the current one
    // (twiddles from LDS, so no vector memory wait inside the transform holds
    // for the prefetch)
    constexpr bool PF = MODE == FWD_STAGE;
    constexpr int TWN = PF ? tw_entries<L, PL>() : 1;
    __shared__ cpx<TW> lds[L * T];
    __shared__ cpx<TW> ltw[TWN];
    const int tid = threadIdx.x;
    const int col0 = (g.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x) * T;
    cpx<TW> v[E];
    TwStage<TWN, NT, TW> tws;
    if constexpr (PF) twiddles_load(tws, tw, tid);
    // kept outputs of the last radix under truncation: m in [0, RP/3) u [2RP/3, RP)
    constexpr bool KEEP = MODE == FWD_STAGE || MODE == FWD_VJP || MODE == FWD_VJPY;
    constexpr int KR = KEEP ? 2 * RP / 3 : 1;
    constexpr int K = KEEP ? (E / RP) * KR : 1;
    static_assert(!KEEP || (RP % 3 == 0 && 3 * LOUT == 2 * L), "stage pass needs a 3/2-padded plan");
    // kept modes per epilogue chunk (their loads are in flight together),
    // sized so the chunk's loads fit beside the kept outputs r without
    // spilling: stages 3 and 4 read three state arrays, the reverse stages
    // read FP64 cotangents in both precisions
    constexpr bool HEAVY = STG == ST_3 || (STG >= RK_1 && STG <= RK_4);
    constexpr int CH = MODE == FWD_VJP ? 1 : sizeof(TW) == 8 ? (HEAVY ? 1 : 2) : (HEAVY ? 2 : 4);
    static_assert(!KEEP || K % CH == 0, "chunks must tile the kept modes");
    cpx<TW> r[KEEP ? 3 : 1][K];

    const int col = col0 + tid % T;
    const bool colok = col < g.ncols && (col % g.hp) < g.hvalid;
    const long long cin = colok ? col_addr(col, g.in_inner, g.in_os) : 0;
    const long long cout = colok ? col_addr(col, g.out_inner, g.out_os) : 0;
    auto load_into = [&](cpx<TW>(&S)[E], int f) {
        const gptr<cpx<TW>> src = opaque(in) + f * g.in_fs + cin;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int p = ci_pos_in<L, PL::first, T, NT>(tid, e);
            S[e] = MODE == FWD_STAGE ? ldg_nt(src, (long long)p * g.in_ls) : ldg(src, (long long)p * g.in_ls);
        }
    };
    auto load_field = [&](int f) { load_into(v, f); };
    // kept element i = b * KR + mk: state-layout index, k^2 (table index; 0 for
    // columns outside the geometry) and integer wavenumbers
    auto kept = [&](int i, long long &idx, int &q, int &wx, int &wy, int &iz) {
        const int b = i / KR, mk = i % KR;
        const int m = mk < RP / 3 ? mk : mk + RP / 3;
        const int p = ci_pos_out<L, RP, T, NT>(tid, b * RP + m);
        const int j = unpad_index(p, LOUT, L);
        const int iy = col / g.hp;
        iz = col % g.hp;
        wx = wavenum(j, n);
        wy = wavenum(iy, n);
        idx = cout + (long long)j * g.out_ls;
        q = colok ? wx * wx + wy * wy + iz * iz : 0;
    };
    // the same for the reverse-stage epilogue, from an opaque copy of the
    // column: the compiler recomputes these integers instead of keeping the
    // curl step's values live across the three GU transforms
    auto kept2 = [&](int i, long long &idx, int &q, int &wx, int &wy, int &iz) {
        int c = col, t = tid;
        asm volatile("" : "+v"(c), "+v"(t));
        const int b = i / KR, mk = i % KR;
        const int m = mk < RP / 3 ? mk : mk + RP / 3;
        const int p = ci_pos_out<L, RP, T, NT>(t, b * RP + m);
        const int j = unpad_index(p, LOUT, L);
        const int iy = c / g.hp;
        iz = c % g.hp;
        wx = wavenum(j, n);
        wy = wavenum(iy, n);
        const bool ok = c < g.ncols && iz < g.hvalid;
        idx = (ok ? col_addr(c, g.out_inner, g.out_os) : 0) + (long long)j * g.out_ls;
        q = ok ? wx * wx + wy * wy + iz * iz : 0;
    };
    // kept output i of the transform in v (scaled)
    auto kept_v = [&](int i) -> cpx<TW> {
        const int b = i / KR, mk = i % KR;
        return cscale(v[b * RP + (mk < RP / 3 ? mk : mk + RP / 3)], scale);
    };
    auto ik = [](TW k, cpx<TW> a) -> cpx<TW> { return {-k * a.im, k * a.re}; };
    if constexpr (MODE == FWD_VJP) {
        // t = GU + curl(GW) (SpectralOps::curl, vonkarman-core/src/spectral_ops.rs:93-104)
        // from the y pass's five fields: 0 t_x (complete), 1 t_y' = GU_y +
        // i kz GW_x, 2 t_z' = GU_z - i ky GW_x, 3 GW_y, 4 GW_z; the i kx terms
        // are added here: t_y = t_y' - i kx GW_z, t_z = t_z' + i kx GW_y
        auto kx_of = [&](int i) -> TW {
            long long idx;
            int q, wx, wy, iz;
            kept(i, idx, q, wx, wy, iz);
            return TW(double(wx) * ks);
        };
        auto field = [&](int f, auto take) {
            load_field(f);
            fft_reg<L, -1, T, NT, T, 1, PL>(v, lds, tw, tid);
#pragma unroll
            for (int i = 0; i < K; ++i) take(i, kept_v(i));
            __syncthreads();
        };
        field(3, [&](int i, cpx<TW> x) { r[2][i] = ik(kx_of(i), x); });
        field(4, [&](int i, cpx<TW> x) { r[1][i] = ik(-kx_of(i), x); });
        field(1, [&](int i, cpx<TW> x) { r[1][i] = cadd(x, r[1][i]); });
        field(2, [&](int i, cpx<TW> x) { r[2][i] = cadd(x, r[2][i]); });
        field(0, [&](int i, cpx<TW> x) { r[0][i] = x; });
#pragma unroll
        for (int c0 = 0; c0 < K; c0 += CH) {
            long long idx[CH];
            int q[CH];
            double kv[CH][3];
            cpx<double> tt[CH][3];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                int wx, wy, iz;
                kept2(c0 + i, idx[i], q[i], wx, wy, iz);
                kv[i][0] = double(wx) * ks;
                kv[i][1] = double(wy) * ks;
                kv[i][2] = double(iz) * ks;
#pragma unroll
                for (int f = 0; f < 3; ++f) tt[i][f] = cvt<double>(r[f][c0 + i]);
            }
            adj_chunk<STG, CH>(sa, colok, idx, q, kv, tt);
        }
    } else if constexpr (MODE == FWD_VJPY) {
        // y forward of the z VJP pass's output (GU: fields 0..2, GW: 3..5),
        // truncated, into the five fields the x VJP pass reads.  The curl
        // terms of t = GU + curl GW along ky and kz are formed here (ky is this
        // pass's wavenumber, kz the column's), the kx terms after the x
        // transform, so one field of traffic and one x transform are saved:
        //   0 t_x  = GU_x + i (ky GW_z - kz GW_y)   (complete)
        //   1 t_y' = GU_y + i kz GW_x
        //   2 t_z' = GU_z - i ky GW_x
        //   3 GW_y, 4 GW_z
        // Padding columns are stored too (into row padding), as in the plain pass.
        constexpr bool PADP = pad_stores<TW, T>();
        const bool stp = PADP || colok;
        const long long coutp = stp ? col_addr(col, g.out_inner, g.out_os) : 0;
        const TW kz = TW(double(col % g.hp) * ks);
        auto jof = [&](int i) -> int {
            const int b = i / KR, mk = i % KR;
            return unpad_index(ci_pos_out<L, RP, T, NT>(tid, b * RP + (mk < RP / 3 ? mk : mk + RP / 3)), LOUT, L);
        };
        auto ky_of = [&](int i) -> TW { return TW(double(wavenum(jof(i), n)) * ks); };
        auto put = [&](int fo, int i, cpx<TW> x) {
            if (stp) out[fo * g.out_fs + coutp + (long long)jof(i) * g.out_ls] = x;
        };
        cpx<TW> h[K];  // GW_x, then i (ky GW_z - kz GW_y)
        auto field = [&](int f, auto take) {
            load_field(f);
            fft_reg<L, -1, T, NT, T, 1, PL>(v, lds, tw, tid);
#pragma unroll
            for (int i = 0; i < K; ++i) take(i, kept_v(i));
            __syncthreads();
        };
        field(3, [&](int i, cpx<TW> x) { h[i] = x; });
        field(1, [&](int i, cpx<TW> x) { put(1, i, cadd(x, ik(kz, h[i]))); });
        field(2, [&](int i, cpx<TW> x) { put(2, i, cadd(x, ik(-ky_of(i), h[i]))); });
        field(4, [&](int i, cpx<TW> x) {
            put(3, i, x);
            h[i] = ik(-kz, x);
        });
        field(5, [&](int i, cpx<TW> x) {
            put(4, i, x);
            h[i] = cadd(h[i], ik(ky_of(i), x));
        });
        field(0, [&](int i, cpx<TW> x) { put(0, i, cadd(x, h[i])); });
    } else if constexpr (MODE == FWD_STAGE) {
        // three fields; the kept outputs of each stay in registers (static indices)
        cpx<TW> Q[E];
        load_field(0);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            if (f + 1 < 3) load_into(Q, f + 1);
            // the LDS twiddle writes follow the first field loads (see twiddles_load)
            if (f == 0) twiddles_store(ltw, tws, tid);
            fft_reg_lt<L, -1, T, NT, T, 1, PL>(v, lds, ltw, tid);
#pragma unroll
            for (int b = 0; b < E / RP; ++b)
#pragma unroll
                for (int mk = 0; mk < KR; ++mk) {
                    const int m = mk < RP / 3 ? mk : mk + RP / 3;
                    r[f][b * KR + mk] = cscale(v[b * RP + m], scale);
                }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = Q[e];
        }
#pragma unroll
        for (int c0 = 0; c0 < K; c0 += CH) {
            long long idx[CH];
            int q[CH];
            double k2d[CH];
            cpx<TW> nn[CH][3];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                int wx, wy, iz;
                kept(c0 + i, idx[i], q[i], wx, wy, iz);
                const double kxd = double(wx) * ks, kyd = double(wy) * ks, kzd = double(iz) * ks;
                k2d[i] = kxd * kxd + kyd * kyd + kzd * kzd;
                const TW kx = TW(kxd), ky = TW(kyd), kz = TW(kzd);
                // Leray projection (vonkarman-core/src/spectral_ops.rs:114-151)
                cpx<TW> x = r[0][c0 + i], y = r[1][c0 + i], z = r[2][c0 + i];
                const TW k2 = kx * kx + ky * ky + kz * kz;
                if (!(double(k2) < 1e-30)) {
                    const TW inv = TW(1) / k2;
                    const TW dr = kx * x.re + ky * y.re + kz * z.re;
                    const TW di = kx * x.im + ky * y.im + kz * z.im;
                    x = {x.re - kx * dr * inv, x.im - kx * di * inv};
                    y = {y.re - ky * dr * inv, y.im - ky * di * inv};
                    z = {z.re - kz * dr * inv, z.im - kz * di * inv};
                }
                nn[i][0] = x;
                nn[i][1] = y;
                nn[i][2] = z;
            }
            stage_chunk<STG, CH, TW>(sa, colok, idx, q, nn, nullptr, k2d);
        }
    } else {
        // Every column stores (padding columns into row padding, see
        // pad_stores), and whether an element's outputs are kept under the
        // truncation is decided at compile time wherever its whole position
        // range is kept or dropped: no store sits under a per-lane branch.
        // The one-field instance (the standalone R2C's y and x passes) stores
        // valid columns only, so its x pass can write the caller's unpadded
        // rows of nz/2+1 directly.
        constexpr bool PADP = pad_stores<TW, T>() && NF != 1;
        const bool stp = PADP || colok;
        const long long coutp = stp ? col_addr(col, g.out_inner, g.out_os) : 0;
        auto store_field = [&](int f) {
            if (stp) {
                static_for<E>([&](auto ec) {
                    constexpr int e = decltype(ec)::value;
                    constexpr int cls = LOUT == L ? 0 : gap_class(ci_out_base<L, RP, T, NT>(e), NT / T, LOUT, L);
                    if constexpr (cls != 1) {
                        const int p = ci_pos_out<L, RP, T, NT>(tid, e);
                        const int j = LOUT == L ? p : unpad_index(p, LOUT, L);
                        if (cls == 0 || j >= 0) out[f * g.out_fs + coutp + (long long)j * g.out_ls] = cscale(v[e], scale);
                    }
                });
            }
        };
        for (int f = 0; f < NF; ++f) {
            load_field(f);
            fft_reg<L, -1, T, NT, T, 1, PL>(v, lds, tw, tid);
            store_field(f);
            __syncthreads();
        }
    }
}

// ===========================================================================
// x forward + truncation + Leray + ETD-RK4 stage epilogue + the next stage's
// x inverse (pad + i kx terms), fused in one pass over the same columns (y, z):
// the stage output (the next stage input) is stored only when sa.out is set.
// The forward plan leaves element e at position j + m L/RP, which is where
// the reversed plan's first pass expects element e, and 3/2 truncation
// followed by zero padding keeps exactly those positions, so the kept stage
// outputs feed the inverse from registers with zeros in the gap elements.
// Outputs (5 fields, the x pass layout of the next nonlinear evaluation):
// 0 ux, 1 uy, 2 uz, 3 i kx uz, 4 i kx uy.
// ===========================================================================
// The fused pass runs the stage plan of its precision in every stage.  FP32
// stage 3 (three state arrays per kept mode) takes one kept mode per epilogue
// chunk so it fits two waves per SIMD: 4-wave blocks, two per CU (the
// 12-element plan needs 226 VGPRs and 8-wave blocks, one per CU: fused pass
// 0.4315 -> 0.4219 ms at 384, same box, the same bits).
template <int L, typename TW, int STG>
struct FusedPlan {
    using type = typename StagePlan<L, TW>::type;
};
template <int L, int STG, typename TW>
using FusedGeom = Geom<L, FusedPlan<L, TW, STG>::type::E, TW, kRowBytes>;

template <int L, int LOUT, int STG, typename TW>
__global__ __launch_bounds__((FusedGeom<L, STG, TW>::NT)) __attribute__((amdgpu_waves_per_eu(Waves<K_FWD, TW>::value))) void k_fwd_stage_inv(
    const cpx<TW> *__restrict__ in, cpx<TW> *__restrict__ aout, const cpx<TW> *__restrict__ tw, Strided g, Strided gi,
    int n, double ks, StageArgs sa) {
    using P = typename FusedPlan<L, TW, STG>::type;
    using PL = typename P::radices;         // forward order
    using RL = typename Reverse<PL>::type;  // inverse order, first radix = forward's last
    constexpr int E = P::E;
    using G = FusedGeom<L, STG, TW>;
    constexpr int T = G::T, NT = G::NT;
    constexpr int RP = PL::last;
    constexpr int KR = 2 * RP / 3, K = (E / RP) * KR;
    static_assert(RP % 3 == 0 && 3 * LOUT == 2 * L, "stage pass needs a 3/2-padded plan");
    constexpr bool HEAVY = STG == ST_3;
    // kept modes per epilogue chunk.  FP64 takes one: two spilled 40 B per
    // lane in stages 1-2 (fused pass 0.663 -> 0.648 ms with one, same box;
    // with the current epilogue two spill 24 B and ran 0.632 -> 0.644 ms).
    // Issuing each next chunk's loads before this chunk's stores (a two-deep
    // pipeline) spilled 80-392 B per lane, as did larger chunks.
    constexpr int CH = sizeof(TW) == 8 || HEAVY ? 1 : 4;
    static_assert(K % CH == 0, "chunks must tile the kept modes");
    constexpr int TWN = tw_entries<L, PL>();
    __shared__ cpx<TW> lds[L * T];
    __shared__ cpx<TW> ltw[TWN];
    const int tid = threadIdx.x;
    const int col0 = (g.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x) * T;
    cpx<TW> v[E], Q[E];
    cpx<TW> r[3][K];
    TwStage<TWN, NT, TW> tws;
    twiddles_load(tws, tw, tid);

    const int col = col0 + tid % T;
    const bool colok = col < g.ncols && (col % g.hp) < g.hvalid;
    const long long cin = colok ? col_addr(col, g.in_inner, g.in_os) : 0;
    const long long cout = colok ? col_addr(col, g.out_inner, g.out_os) : 0;
    constexpr bool PADST = pad_stores<TW, T>();
    const long long cinv = (PADST || colok) ? col_addr(col, gi.out_inner, gi.out_os) : 0;
    auto load_into = [&](cpx<TW>(&S)[E], int f) {
        const gptr<cpx<TW>> src = opaque(in) + f * g.in_fs + cin;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const long long i = (long long)ci_pos_in<L, PL::first, T, NT>(tid, e) * g.in_ls;
            S[e] = ldg_nt(src, i);
        }
    };
    // forward transforms of the three fields (each next field loads meanwhile)
    load_into(v, 0);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        if (f + 1 < 3) load_into(Q, f + 1);
        // the LDS twiddle writes follow the first field loads (see twiddles_load)
        if (f == 0) twiddles_store(ltw, tws, tid);
        fft_reg_lt<L, -1, T, NT, T, 1, PL>(v, lds, ltw, tid);
#pragma unroll
        for (int b = 0; b < E / RP; ++b)
#pragma unroll
            for (int mk = 0; mk < KR; ++mk) r[f][b * KR + mk] = v[b * RP + (mk < RP / 3 ? mk : mk + RP / 3)];
        __syncthreads();
        if (f + 1 < 3) {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = Q[e];
        }
    }
    // Leray + stage epilogue per kept mode; r <- the stage output
#pragma unroll
    for (int c0 = 0; c0 < K; c0 += CH) {
        long long idx[CH];
        int q[CH];
        cpx<TW> nn[CH][3], xo[CH][3];
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int b = (c0 + i) / KR, mk = (c0 + i) % KR;
            const int p = ci_pos_out<L, RP, T, NT>(tid, b * RP + (mk < RP / 3 ? mk : mk + RP / 3));
            const int j = unpad_index(p, LOUT, L);
            const int iy = col / g.hp, iz = col % g.hp;
            const int wx = wavenum(j, n), wy = wavenum(iy, n);
            idx[i] = cout + (long long)j * g.out_ls;
            q[i] = colok ? wx * wx + wy * wy + iz * iz : 0;
            const TW kx = TW(double(wx) * ks), ky = TW(double(wy) * ks), kz = TW(double(iz) * ks);
            // Leray projection (vonkarman-core/src/spectral_ops.rs:114-151)
            cpx<TW> x = r[0][c0 + i], y = r[1][c0 + i], z = r[2][c0 + i];
            const TW k2 = kx * kx + ky * ky + kz * kz;
            if (!(double(k2) < 1e-30)) {
                const TW inv = TW(1) / k2;
                const TW dr = kx * x.re + ky * y.re + kz * z.re;
                const TW di = kx * x.im + ky * y.im + kz * z.im;
                x = {x.re - kx * dr * inv, x.im - kx * di * inv};
                y = {y.re - ky * dr * inv, y.im - ky * di * inv};
                z = {z.re - kz * dr * inv, z.im - kz * di * inv};
            }
            nn[i][0] = x;
            nn[i][1] = y;
            nn[i][2] = z;
        }
        stage_chunk<STG, CH, TW, true>(sa, colok, idx, q, nn, xo);
#pragma unroll
        for (int i = 0; i < CH; ++i)
#pragma unroll
            for (int f = 0; f < 3; ++f) r[f][c0 + i] = xo[i][f];
    }
    // x inverse of the stage output: element e of the reversed plan's input is
    // forward output element e (zero in the truncation gap)
    auto xin = [&](int f, int e) -> cpx<TW> {
        const int b = e / RP, m = e % RP;
        if (m >= RP / 3 && m < 2 * RP / 3) return {TW(0), TW(0)};
        return r[f][b * KR + (m < RP / 3 ? m : m - RP / 3)];
    };
    auto kx_of = [&](int e) -> TW {
        const int j = unpad_index(ci_pos_out<L, RP, T, NT>(tid, e), LOUT, L);
        return TW(double(wavenum(j < 0 ? 0 : j, n)) * ks);
    };
    auto ik = [](TW k, cpx<TW> a) -> cpx<TW> { return {-k * a.im, k * a.re}; };
    auto emit = [&](int fo, auto value) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = value(e);
        fft_reg_lt<L, +1, T, NT, T, 1, RL>(v, lds, ltw, tid);
        if (PADST || colok) {
            // the store rows from a fresh copy of the thread index: the
            // compiler otherwise keeps them from the load phase through the
            // whole pass (stage 3 spilled 96 B per lane in FP64, 64 B in FP32;
            // 72 B and 20 B with this, the same time)
            const int t = pinned(tid);
#pragma unroll
            for (int e = 0; e < E; ++e)
            {
                cpx<TW> *dst = aout + fo * gi.out_fs + cinv + (long long)ci_pos_out<L, RL::last, T, NT>(t, e) * gi.out_ls;
                stg_nt(dst, v[e]);
            }
        }
        __syncthreads();
    };
    emit(0, [&](int e) { return xin(0, e); });
    emit(1, [&](int e) { return xin(1, e); });
    emit(4, [&](int e) { return ik(kx_of(e), xin(1, e)); });
    emit(2, [&](int e) { return xin(2, e); });
    emit(3, [&](int e) { return ik(kx_of(e), xin(2, e)); });
}

template <int L, int LOUT, int STG, typename TW>
int run_fwd_inv(const void *in, void *aout, const void *tw, const Strided &g, const Strided &gi, int n, double ks,
                const StageArgs *sa, hipStream_t st) {
    using C = FusedGeom<L, STG, TW>;
    if (pad_stores<TW, C::T>() && g.ncols % C::T != 0) return 2;
    const unsigned nb = (unsigned)((g.ncols + C::T - 1) / C::T);
    hipLaunchKernelGGL((k_fwd_stage_inv<L, LOUT, STG, TW>), dim3(nb), dim3(C::NT), 0, st,
                       reinterpret_cast<const cpx<TW> *>(in), reinterpret_cast<cpx<TW> *>(aout),
                       reinterpret_cast<const cpx<TW> *>(tw), g, gi, n, ks, *sa);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

template <int L>
int launch_fwd_inv_L(int lout, Prec wp, const void *in, void *aout, const void *tw, const Strided &g,
                     const Strided &gi, int n, double ks, const StageArgs *sa, hipStream_t st) {
    if constexpr (L % 3 == 0 && (2 * L / 3) % 2 == 0) {
        constexpr int LOUT = 2 * L / 3;
        if (lout != LOUT || !sa) return 2;
        switch (sa->stage) {
            case ST_1:
                return wp == F64 ? run_fwd_inv<L, LOUT, ST_1, double>(in, aout, tw, g, gi, n, ks, sa, st)
                                 : run_fwd_inv<L, LOUT, ST_1, float>(in, aout, tw, g, gi, n, ks, sa, st);
            case ST_2:
                return wp == F64 ? run_fwd_inv<L, LOUT, ST_2, double>(in, aout, tw, g, gi, n, ks, sa, st)
                                 : run_fwd_inv<L, LOUT, ST_2, float>(in, aout, tw, g, gi, n, ks, sa, st);
            case ST_3:
                return wp == F64 ? run_fwd_inv<L, LOUT, ST_3, double>(in, aout, tw, g, gi, n, ks, sa, st)
                                 : run_fwd_inv<L, LOUT, ST_3, float>(in, aout, tw, g, gi, n, ks, sa, st);
            case ST_4:
                return wp == F64 ? run_fwd_inv<L, LOUT, ST_4, double>(in, aout, tw, g, gi, n, ks, sa, st)
                                 : run_fwd_inv<L, LOUT, ST_4, float>(in, aout, tw, g, gi, n, ks, sa, st);
        }
    }
    return 2;
}

// ===========================================================================
// Contiguous (z) passes.  Two real lines travel through one complex DFT:
//   C2R:  Z_k = X_k + i Y_k,  Z_{L-k} = conj(X_k) + i conj(Y_k)  ->  x + i y
//   R2C:  A_k = (Z_k + conj Z_{L-k})/2,  B_k = (Z_k - conj Z_{L-k})/(2i)
// The imaginary parts of the k=0 and k=L/2 inputs are ignored by the C2R
// (vonkarman-fft/src/backend.rs:15-18 semantics).  Each block holds T lines
// (line-major tile, pitch L); registers are filled straight from HBM.
// ===========================================================================
// Hermitian-extended pair spectrum Z at FFT position p from the half
// spectra fa, fb (HIN stored entries; fb absent when HASB is false):
//   p < HIN (stored):        Z = X_p + i Y_p      (Im X, Im Y dropped at p = 0, L/2)
//   p > L - HIN (mirrored):  Z = conj X_{L-p} + i conj Y_{L-p}
//   otherwise (zero gap):    Z = 0
// CLS classifies the whole position range of the load at compile time
// (zpair_class); in the mixed classes the loads are unconditional from a
// clamped index and the cases are selected afterwards, because a load under a
// branch is waited for before the branch joins.
//   ZP_DIRECT0:  stored, the edge p = 0 included (range starting at 0)
//   ZP_DIRZERO:  stored below HIN, zero gap above (no edge, no mirror)
//   ZP_ZEROMIR:  zero gap up to L - HIN, mirrored above (no edge)
enum { ZP_DIRECT = 0, ZP_MIRROR = 1, ZP_ZERO = 2, ZP_MIXED = 3, ZP_DIRECT0 = 4, ZP_DIRZERO = 5, ZP_ZEROMIR = 6 };
template <int L, int HIN>
constexpr int zpair_class(int b, int w) {
    const int e = b + w;  // exclusive end of the range
    const bool half = b <= L / 2 && L / 2 < e;
    const bool edge = b == 0 || half;
    if (!edge && b >= 1 && e <= HIN) return ZP_DIRECT;
    if (!edge && b >= HIN && b > L - HIN && e <= L) return ZP_MIRROR;
    if (b >= HIN && e - 1 <= L - HIN) return ZP_ZERO;
    if (b == 0 && !half && e <= HIN) return ZP_DIRECT0;
    if (!edge && b >= 1 && b < HIN && e - 1 <= L - HIN) return ZP_DIRZERO;
    if (!edge && b >= HIN && e <= L) return ZP_ZEROMIR;
    return ZP_MIXED;
}
// One element of the pair spectrum in two steps: the raw loads (X from fa, Y
// from fb; nothing for the zero class) and their combination into Z, so the
// next pair's loads can be issued before the current pair's transform.
template <typename TW>
struct ZRaw {
    cpx<TW> x, y;
};
template <int L, int HIN, bool HASB, int CLS, typename TW>
__device__ __forceinline__ void z_pair_raw(ZRaw<TW> &r, gptr<cpx<TW>> fa, gptr<cpx<TW>> fb, long long base, int p) {
    if constexpr (CLS != ZP_ZERO) {
        int k;
        if constexpr (CLS == ZP_DIRECT || CLS == ZP_DIRECT0) k = p;
        else if constexpr (CLS == ZP_MIRROR) k = L - p;
        else if constexpr (CLS == ZP_DIRZERO) k = p < HIN ? p : HIN - 1;
        else if constexpr (CLS == ZP_ZEROMIR) k = p > L - HIN ? L - p : HIN - 1;
        else k = p < HIN ? p : (p > L - HIN ? L - p : 0);
        r.x = ldg(fa, base + k);
        if constexpr (HASB) r.y = ldg(fb, base + k);
    }
}
template <int L, int HIN, bool HASB, int CLS, typename TW>
__device__ __forceinline__ cpx<TW> z_pair_make(const ZRaw<TW> &r, int p) {
    const cpx<TW> zero = {TW(0), TW(0)};
    if constexpr (CLS == ZP_ZERO) {
        return zero;
    } else {
        const cpx<TW> X = r.x;
        const cpx<TW> Y = HASB ? r.y : zero;
        const cpx<TW> dir = {X.re - Y.im, X.im + Y.re}, mir = {X.re + Y.im, Y.re - X.im};
        if constexpr (CLS == ZP_DIRECT) {
            return dir;
        } else if constexpr (CLS == ZP_MIRROR) {
            return mir;
        } else if constexpr (CLS == ZP_DIRECT0) {
            // at p = 0 the imaginary parts are dropped: Z = Re X + i Re Y
            return p == 0 ? cpx<TW>{X.re, Y.re} : dir;
        } else if constexpr (CLS == ZP_DIRZERO) {
            return p < HIN ? dir : zero;
        } else if constexpr (CLS == ZP_ZEROMIR) {
            return p > L - HIN ? mir : zero;
        } else {
            const bool lo = p < HIN, hi = !lo && p > L - HIN, edge = p == 0 || 2 * p == L;
            const TW xi = edge ? TW(0) : X.im, yi = edge ? TW(0) : Y.im;
            const TW zr = lo ? X.re - yi : (hi ? X.re + yi : TW(0));
            const TW zi = lo ? xi + Y.re : (hi ? Y.re - xi : TW(0));
            return {zr, zi};
        }
    }
}
// the raw loads of a column-interleaved z tile (T lines, NT threads, first
// radix R1) for the pair spectrum of line `base` (element e at position
// ci_pos_in(tid, e)), and their combination into the tile's registers
template <int L, int HIN, int R1, int T, int NT, bool HASB, int E, typename TW>
__device__ __forceinline__ void z_pair_issue(ZRaw<TW> (&r)[E], gptr<cpx<TW>> fa, gptr<cpx<TW>> fb, long long base,
                                             int tid) {
    static_for<E>([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        constexpr int cls = zpair_class<L, HIN>(ci_in_base<L, R1, T, NT>(e), NT / T);
        z_pair_raw<L, HIN, HASB, cls, TW>(r[e], fa, fb, base, ci_pos_in<L, R1, T, NT>(tid, e));
    });
}
template <int L, int HIN, int R1, int T, int NT, bool HASB, int E, typename TW>
__device__ __forceinline__ void z_pair_combine(cpx<TW> (&v)[E], const ZRaw<TW> (&r)[E], int tid) {
    static_for<E>([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        constexpr int cls = zpair_class<L, HIN>(ci_in_base<L, R1, T, NT>(e), NT / T);
        v[e] = z_pair_make<L, HIN, HASB, cls, TW>(r[e], ci_pos_in<L, R1, T, NT>(tid, e));
    });
}
template <int L, int HIN, int R1, int T, int NT, bool HASB, int E, typename TW>
__device__ __forceinline__ void z_pair_fill(cpx<TW> (&v)[E], gptr<cpx<TW>> fa, gptr<cpx<TW>> fb, long long base,
                                            int tid) {
    ZRaw<TW> r[E];
    z_pair_issue<L, HIN, R1, T, NT, HASB>(r, fa, fb, base, tid);
    z_pair_combine<L, HIN, R1, T, NT, HASB>(v, r, tid);
}

// z tiles are column-interleaved too: the T lines of a block are the columns,
// so position p of all T lines is one contiguous LDS row.  Two lines per block
// make a block of one wavefront (E = 12, NT = 64): the compiler drops its
// barriers, and blocks on a SIMD run independently.  Each line still reads
// HBM in runs of 32 consecutive positions per load instruction.  Measured at
// 384 (256^3 solver), per launch: FP64 1.03 ms with 4 lines per block -> 0.87
// ms with 2; FP32 0.65 ms with 4 -> 0.50 ms with 2.
//
// z passes read their twiddles from a per-block LDS copy: a global twiddle
// load inside a radix pass is a dependent memory wait that a one-wave block
// cannot hide.
template <int L, int D, int T, int NT, typename PL, int E, typename R>
__device__ __forceinline__ void zfft(cpx<R> (&v)[E], cpx<R> *lds, const cpx<R> *ltw, int tid) {
    fft_reg_lt<L, D, T, NT, T, 1, PL>(v, lds, ltw, tid);
}
template <int L, typename TW>
using ZNLP = PlanE12<L>;
// T = 2 lines per block: rows of 32 B (FP64) or 16 B (FP32)
template <int L, typename TW>
using ZNLGeom = Geom<L, ZNLP<L, TW>::E, TW, 2 * 2 * (int)sizeof(TW)>;

template <int L, typename TW>
__global__ __launch_bounds__((ZNLGeom<L, TW>::NT)) __attribute__((amdgpu_waves_per_eu(Waves<K_ZNL, TW>::value))) void k_z_nl(const cpx<TW> *__restrict__ in,
                                                               cpx<TW> *__restrict__ out, long long lines, int hp,
                                                               long long fs_in, long long fs_out, TW s_in, TW s_out,
                                                               const cpx<TW> *__restrict__ tw) {
    using P = ZNLP<L, TW>;
    using PL = typename P::radices;         // inverse order
    using RL = typename Reverse<PL>::type;  // forward order: first radix = inverse's last
    constexpr int E = P::E;
    constexpr int T = ZNLGeom<L, TW>::T, NT = ZNLGeom<L, TW>::NT;
    constexpr int HIN = L / 3 + 1;  // N/2+1 with N = 2L/3
    constexpr int TWN = tw_entries<L, PL>();
    __shared__ cpx<TW> lds[L * T];
    __shared__ cpx<TW> ltw[TWN];
    const int tid = threadIdx.x;
    const long long line0 = (long long)blockIdx.x * T;
    TwStage<TWN, NT, TW> tws;
    twiddles_load(tws, tw, tid);
    const int t = tid % T;
    const bool ok = line0 + t < lines;
    // lines past the end load the last line (valid memory); their results are not stored
    const long long base = (ok ? line0 + t : lines - 1) * hp;
    cpx<TW> v[E];
    TW r[6][E];

    // fields: 0 ux, 1 uy, 2 uz, 3 wx, 4 wy, 5 wz; pairs (0,1) (2,3) (4,5).
    // Each next pair's loads are in flight during the current pair's
    // transform (at 384: FP64 0.738 -> 0.704 ms; FP32 0.395 -> 0.387 ms, at
    // 128 VGPRs under its 3-wave floor).
    ZRaw<TW> raw[E];
    auto issue = [&](int q) {
        z_pair_issue<L, HIN, PL::first, T, NT, true>(raw, opaque(in) + (2 * q) * fs_in,
                                                      opaque(in) + (2 * q + 1) * fs_in, base, tid);
    };
    issue(0);
    // the LDS twiddle writes follow the first pair's loads (see twiddles_load)
    twiddles_store(ltw, tws, tid);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        z_pair_combine<L, HIN, PL::first, T, NT, true>(v, raw, tid);
        if (q + 1 < 3) issue(q + 1);
        zfft<L, +1, T, NT, PL>(v, lds, ltw, tid);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            r[2 * q][e] = v[e].re * s_in;
            r[2 * q + 1][e] = v[e].im * s_in;
        }
        __syncthreads();
    }
    // u x w (vonkarman-fft/src/dealiased.rs:79-81)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const TW ux = r[0][e], uy = r[1][e], uz = r[2][e];
        const TW ox = r[3][e], oy = r[4][e], oz = r[5][e];
        r[0][e] = uy * oz - uz * oy;
        r[1][e] = uz * ox - ux * oz;
        r[2][e] = ux * oy - uy * ox;
    }
    constexpr int HOUT = HIN;
    // (c0 + i c1): forward, then split through LDS (needs k and L-k)
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = {r[0][e], r[1][e]};
    zfft<L, -1, T, NT, RL>(v, lds, ltw, tid);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) lds[ci_pos_out<L, RL::last, T, NT>(tid, e) * T + t] = v[e];
    __syncthreads();
    const TW h = TW(0.5) * s_out;
    // Output lines are written whole, entries [HOUT, hp) as zeros: a line
    // written only up to HOUT leaves its last cache line partly written,
    // which costs HBM time (no pass reads those entries).
    const cpx<TW> zero = {TW(0), TW(0)};
    for (int e = tid; e < T * hp; e += NT) {
        const int tt = e % T, k = e / T;
        const long long line = line0 + tt;
        const bool kept = k < HOUT;
        const int kk = kept ? k : 0;
        const cpx<TW> zk = lds[kk * T + tt], zm = lds[(kk ? L - kk : 0) * T + tt];
        if (line < lines) {
            out[0 * fs_out + line * hp + k] = kept ? cpx<TW>{(zk.re + zm.re) * h, (zk.im - zm.im) * h} : zero;
            out[1 * fs_out + line * hp + k] = kept ? cpx<TW>{(zk.im + zm.im) * h, (zm.re - zk.re) * h} : zero;
        }
    }
    __syncthreads();
    // c2 alone: its first hp outputs are stored from registers
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = {r[2][e], TW(0)};
    zfft<L, -1, T, NT, RL>(v, lds, ltw, tid);
    if (ok) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int p = ci_pos_out<L, RL::last, T, NT>(tid, e);
            if (p < hp) out[2 * fs_out + base + p] = p < HOUT ? cscale(v[e], s_out) : zero;
        }
    }
}

// ===========================================================================
// The z pass at L = 384 with one line per wavefront.  384 = 6 * 8 * 8, and a
// transform moves a line between two register layouts:
//   layout 64: lane b (all 64 lanes) holds positions 64 a + b, a = 0..5;
//   layout 48: lane j = c + 6 d0 (48 lanes) holds positions j + 48 d1 =
//   c + 6 d0 + 48 d1, d1 = 0..7, so its loads and stores are coalesced.
//   Lanes 48..63 repeat lanes 0..15 and write nothing.
// zw_64to48 is the DFT of direction D as
//   radix-6 over a, times w_384^{D b c}; exchange; radix-8 over b1 (with
//   b = 8 b1 + b0; lane 8 c + b0), times w_64^{D b0 d0}; exchange; radix-8
//   over b0,
// and zw_48to64 the same passes in reverse order, each twiddle applied before
// its butterfly (the DFT matrix is symmetric, so both compute the DFT of
// direction D, from one layout into the other).  Which side takes which
// layout is a template choice (S64): FP64 puts the spectral side in layout 48
// and the physical side in layout 64, so the six physical fields of the cross
// product take 6 elements per lane; FP32 puts the spectral side in layout 64
// (every lane loads and stores).  A lane holds at most four fields plus one
// transform at once, against 12 elements per field when 32 lanes share a
// line, so the pass runs at more waves per SIMD; each transform makes two LDS
// exchanges instead of three.
// ===========================================================================
// index maps of the two layouts and the exchanges: zw:: in vk_fft.hpp

// v[m] *= w^m (D < 0) or conj(w)^m (D > 0), m = 1..N-1.  The powers are
// formed anew in every transform: w1 passes through an empty asm statement,
// so the compiler cannot keep the powers of all transforms in registers.
template <int D, int N, typename R>
__device__ __forceinline__ void zw_twiddle(cpx<R> (&v)[N], cpx<R> w1) {
    asm volatile("" : "+v"(w1.re), "+v"(w1.im));
    cpx<R> w = w1;
#pragma unroll
    for (int m = 1; m < N; ++m) {
        v[m] = D < 0 ? cmul(v[m], w) : cmulc(v[m], w);
        if (m + 1 < N) w = cmul(w, w1);
    }
}

// The twiddles of a lane, wa = W_384^b (lane b) and wb = W_64^{j % 8}: the
// bases only, their powers formed in every transform (TC = false), or all
// twelve powers formed once (TC = true: 10 complex registers more, 10 complex
// products fewer per transform).
template <typename R, bool TC>
struct ZwTw;
template <typename R>
struct ZwTw<R, false> {
    cpx<R> wa, wb;
    __device__ __forceinline__ ZwTw(cpx<R> a, cpx<R> b) : wa(a), wb(b) {}
    template <int D>
    __device__ __forceinline__ void a6(cpx<R> (&v)[6]) const { zw_twiddle<D>(v, wa); }
    template <int D>
    __device__ __forceinline__ void b8(cpx<R> (&v)[8]) const { zw_twiddle<D>(v, wb); }
};
template <typename R>
struct ZwTw<R, true> {
    cpx<R> pa[5], pb[7];  // w^1..w^5 of wa, w^1..w^7 of wb
    __device__ __forceinline__ ZwTw(cpx<R> a, cpx<R> b) {
        pa[0] = a;
        pb[0] = b;
#pragma unroll
        for (int m = 1; m < 5; ++m) pa[m] = cmul(pa[m - 1], a);
#pragma unroll
        for (int m = 1; m < 7; ++m) pb[m] = cmul(pb[m - 1], b);
    }
    template <int D, int N>
    __device__ __forceinline__ static void ap(cpx<R> (&v)[N], const cpx<R> *p) {
#pragma unroll
        for (int m = 1; m < N; ++m) v[m] = D < 0 ? cmul(v[m], p[m - 1]) : cmulc(v[m], p[m - 1]);
    }
    template <int D>
    __device__ __forceinline__ void a6(cpx<R> (&v)[6]) const { ap<D>(v, pa); }
    template <int D>
    __device__ __forceinline__ void b8(cpx<R> (&v)[8]) const { ap<D>(v, pb); }
};

// tw: the lane's twiddles (ZwTw).  The caller synchronises before its next
// LDS write.
template <int D, typename R, typename TWS>
__device__ __forceinline__ void zw_64to48(cpx<R> (&v6)[6], cpx<R> (&v8)[8], cpx<R> *lds, int lane, int j,
                                          const TWS &tw) {
    dft<6, D>(v6);
    tw.template a6<D>(v6);
#pragma unroll
    for (int c = 0; c < 6; ++c) lds[zw::x1(c, lane)] = v6[c];
    __syncthreads();
    {
        const int c = j >> 3, b0 = j & 7;
#pragma unroll
        for (int b1 = 0; b1 < 8; ++b1) v8[b1] = lds[zw::x1(c, 8 * b1 + b0)];
        dft<8, D>(v8);
        tw.template b8<D>(v8);
        __syncthreads();
        if (lane < 48) {
#pragma unroll
            for (int d0 = 0; d0 < 8; ++d0) lds[zw::x2<sizeof(cpx<R>)>(c, d0, b0)] = v8[d0];
        }
    }
    __syncthreads();
    {
        const int c = j % 6, d0 = j / 6;
#pragma unroll
        for (int b0 = 0; b0 < 8; ++b0) v8[b0] = lds[zw::x2<sizeof(cpx<R>)>(c, d0, b0)];
        dft<8, D>(v8);
    }
}

template <int D, typename R, typename TWS>
__device__ __forceinline__ void zw_48to64(cpx<R> (&v8)[8], cpx<R> (&v6)[6], cpx<R> *lds, int lane, int j,
                                          const TWS &tw) {
    dft<8, D>(v8);
    if (lane < 48) {
        const int c = j % 6, d0 = j / 6;
#pragma unroll
        for (int b0 = 0; b0 < 8; ++b0) lds[zw::x2<sizeof(cpx<R>)>(c, d0, b0)] = v8[b0];
    }
    __syncthreads();
    {
        const int c = j >> 3, b0 = j & 7;
#pragma unroll
        for (int d0 = 0; d0 < 8; ++d0) v8[d0] = lds[zw::x2<sizeof(cpx<R>)>(c, d0, b0)];
        tw.template b8<D>(v8);
        dft<8, D>(v8);
        __syncthreads();
        if (lane < 48) {
#pragma unroll
            for (int b1 = 0; b1 < 8; ++b1) lds[zw::x1(c, 8 * b1 + b0)] = v8[b1];
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 6; ++c) v6[c] = lds[zw::x1(c, lane)];
    tw.template a6<D>(v6);
    dft<6, D>(v6);
}

// an empty asm statement that reads v: code after it runs after v is formed
template <typename R, int N>
__device__ __forceinline__ void zw_after(const cpx<R> (&v)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) asm volatile("" ::"v"(v[i].re), "v"(v[i].im));
}

// pair spectrum of one line (fields fa, fb) in layout 64 (S64) or layout 48:
// raw loads and their combination (z_pair_raw / z_pair_make per slot class)
template <bool S64>
struct ZwSpec {
    static constexpr int NS = S64 ? 6 : 8, W = S64 ? 64 : 48;
    static __device__ __forceinline__ int pos(int lane, int j, int s) { return S64 ? 64 * s + lane : zw::pos48(j, s); }
};
template <bool S64, typename TW, bool HASB = true>
__device__ __forceinline__ void zw_issue(ZRaw<TW> (&r)[ZwSpec<S64>::NS], gptr<cpx<TW>> fa, gptr<cpx<TW>> fb,
                                         long long base, int lane, int j) {
    using SP = ZwSpec<S64>;
    static_for<SP::NS>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int cls = zpair_class<zw::L, zw::HIN>(SP::W * s, SP::W);
        z_pair_raw<zw::L, zw::HIN, HASB, cls, TW>(r[s], fa, fb, base, SP::pos(lane, j, s));
    });
}
template <bool S64, typename TW, bool HASB = true>
__device__ __forceinline__ void zw_combine(cpx<TW> (&v)[ZwSpec<S64>::NS], const ZRaw<TW> (&r)[ZwSpec<S64>::NS],
                                           int lane, int j) {
    using SP = ZwSpec<S64>;
    static_for<SP::NS>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int cls = zpair_class<zw::L, zw::HIN>(SP::W * s, SP::W);
        v[s] = z_pair_make<zw::L, zw::HIN, HASB, cls, TW>(r[s], SP::pos(lane, j, s));
    });
}

// S64: the spectral side in layout 64 (physical in layout 48), else the
// reverse.  PF: each next pair's loads in flight during the current pair's
// transform.  WPE: the waves-per-SIMD floor asked of the compiler.  LPB lines
// per block (1 or 2): with two, the c2 of the first line waits in LDS and the
// two c2 share one forward transform (9 transforms per two lines, not 10).
template <typename TW, bool S64, bool PF, int WPE, int LPB, bool TC = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_z_nl_w(
    const cpx<TW> *__restrict__ in, cpx<TW> *__restrict__ out, long long lines, int hp, long long fs_in,
    long long fs_out, TW s_in, TW s_out, const cpx<TW> *__restrict__ tw) {
    using SP = ZwSpec<S64>;
    constexpr int NS = SP::NS, NP = S64 ? 8 : 6;  // spectral and physical slots
    constexpr int HOUT = zw::HIN, L = zw::L;
    static_assert(LPB == 1 || LPB == 2, "one or two lines per block");
    __shared__ cpx<TW> lds[zw::NLDS];
    __shared__ TW park[LPB == 2 ? NP * 64 : 1];
    const int lane = threadIdx.x;
    const int j = lane < 48 ? lane : lane - 48;
    const bool wr = S64 || lane < 48;  // lanes that own spectral slots
    const gptr<cpx<TW>> gtw = (gptr<cpx<TW>>)tw;
    const ZwTw<TW, TC> tws(ldg(gtw, lane), ldg(gtw, 6 * (j & 7)));
    const cpx<TW> zero = {TW(0), TW(0)};
    // The input scale s_in is applied at the output: (s u) x (s w) = s^2 (u x w)
    // and the forward transform is linear, so the physical fields stay
    // unscaled and the kept modes take s_out s_in^2 (exactly the same values
    // when s_in is a power of two, as 1/N^3 is for the power-of-two grids).
    const TW so = s_out * (s_in * s_in);
    const TW h = TW(0.5) * so;
    cpx<TW> vs[NS], vp[NP];
    auto forward = [&]() {
        if constexpr (S64) zw_48to64<-1>(vp, vs, lds, lane, j, tws);
        else zw_64to48<-1>(vp, vs, lds, lane, j, tws);
    };
    // spectral slots holding kept modes: k < 192 (S64) or k < 144
    constexpr int NK = 3;
    // Forward DFT of (a + i b) in vp, split into the half spectra of a and b
    // (stored at oa, ob).  Z_{L-k} of a kept k >= 1 lies in [L - 191, L - 1]:
    // the last two slots of layout 64 (index p - 256) or the last three of
    // layout 48 (index p - 240), exchanged through LDS; Z_{L-0} = Z_0 is k's
    // own.  Lines are written whole, entries [HOUT, hp) as zeros.
    constexpr int PLO = S64 ? 256 : 240;
    auto forward_split = [&](cpx<TW> *oa, cpx<TW> *ob) {
        forward();
        __syncthreads();
        if (wr) {
#pragma unroll
            for (int s = NS - (S64 ? 2 : 3); s < NS; ++s) lds[SP::pos(lane, j, s) - PLO] = vs[s];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NK; ++s) {
            const int k = SP::pos(lane, j, s);
            // k >= 1 pairs with L - k >= PLO when k <= L - PLO; larger k are
            // not kept and read entry 0
            const int li = (k >= 1 && k <= L - PLO) ? L - k - PLO : 0;
            const cpx<TW> zk = vs[s], zl = lds[li];
            const cpx<TW> zm = k ? zl : zk;
            const bool kept = k < HOUT;
            const cpx<TW> va = kept ? cpx<TW>{(zk.re + zm.re) * h, (zk.im - zm.im) * h} : zero;
            const cpx<TW> vb = kept ? cpx<TW>{(zk.im + zm.im) * h, (zm.re - zk.re) * h} : zero;
            if (wr && (s < NK - 1 || k < hp)) {
                oa[k] = va;
                ob[k] = vb;
            }
        }
        __syncthreads();
    };
    ZRaw<TW> raw[NS];
    // grid = lines / LPB (the launcher checks that LPB divides lines)
    auto issue = [&](int l, int q) {
        const long long base = ((long long)blockIdx.x * LPB + l) * hp;
        zw_issue<S64, TW>(raw, opaque(in) + (2 * q) * fs_in, opaque(in) + (2 * q + 1) * fs_in, base, lane, j);
    };
#pragma unroll
    for (int l = 0; l < LPB; ++l) {
        const long long base = ((long long)blockIdx.x * LPB + l) * hp;
        if (PF) issue(l, 0);
        TW r[6][NP];
        // fields: 0 ux, 1 uy, 2 uz, 3 wx, 4 wy, 5 wz; pairs (0,1) (2,3) (4,5)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (!PF) issue(l, q);
            zw_combine<S64, TW>(vs, raw, lane, j);
            if (PF && q + 1 < 3) {
                // the next pair's loads are issued only once this pair's raw
                // loads are combined (half their registers): an empty asm
                // statement that reads the combined values comes first, and
                // the address of each load passes through a later one (opaque)
                zw_after(vs);
                issue(l, q + 1);
            }
            if constexpr (S64) zw_64to48<+1>(vs, vp, lds, lane, j, tws);
            else zw_48to64<+1>(vs, vp, lds, lane, j, tws);
#pragma unroll
            for (int e = 0; e < NP; ++e) {
                r[2 * q][e] = vp[e].re;
                r[2 * q + 1][e] = vp[e].im;
            }
            __syncthreads();
        }
        // u x w (vonkarman-fft/src/dealiased.rs:79-81)
#pragma unroll
        for (int e = 0; e < NP; ++e) {
            const TW ux = r[0][e], uy = r[1][e], uz = r[2][e];
            const TW ox = r[3][e], oy = r[4][e], oz = r[5][e];
            r[0][e] = uy * oz - uz * oy;
            r[1][e] = uz * ox - ux * oz;
            r[2][e] = ux * oy - uy * ox;
        }
        if (LPB == 2 && l == 0) {
            // c2 of the first line waits in LDS for the second line's
#pragma unroll
            for (int e = 0; e < NP; ++e) park[e * 64 + lane] = r[2][e];
        }
        // (c0 + i c1)
#pragma unroll
        for (int e = 0; e < NP; ++e) vp[e] = {r[0][e], r[1][e]};
        forward_split(out + base, out + fs_out + base);
        if constexpr (LPB == 2) {
            if (l == 1) {
                // (c2 of the first line) + i (c2 of this line)
#pragma unroll
                for (int e = 0; e < NP; ++e) vp[e] = {park[e * 64 + lane], r[2][e]};
                forward_split(out + 2 * fs_out + base - hp, out + 2 * fs_out + base);
            }
        } else {
            // c2 alone: its kept outputs are stored from registers
#pragma unroll
            for (int e = 0; e < NP; ++e) vp[e] = {r[2][e], TW(0)};
            forward();
            cpx<TW> *o2 = out + 2 * fs_out + base;
#pragma unroll
            for (int s = 0; s < NK; ++s) {
                const int k = SP::pos(lane, j, s);
                if (wr && (s < NK - 1 || k < hp)) o2[k] = k < HOUT ? cscale(vs[s], so) : zero;
            }
        }
    }
}

// split the forward DFT of (a + i b) held in registers into the truncated
// half spectra of a and b, times c_k s (c_0 = 1, c_k = 2 otherwise): the
// transpose of the unnormalised C2R applied to a and b.  Lines are written
// whole, entries [HOUT, hp) as zeros.
template <int L, int T, int NT, typename RL, int E, int HOUT, typename TW>
__device__ __forceinline__ void z_split_store(const cpx<TW> (&v)[E], cpx<TW> *lds, int tid, long long line0,
                                              long long lines, int hp, cpx<TW> *oa, cpx<TW> *ob, TW s) {
    const int t = tid % T;
    const cpx<TW> zero = {TW(0), TW(0)};
#pragma unroll
    for (int e = 0; e < E; ++e) lds[ci_pos_out<L, RL::last, T, NT>(tid, e) * T + t] = v[e];
    __syncthreads();
    for (int e = tid; e < T * hp; e += NT) {
        const int tt = e % T, k = e / T;
        const long long line = line0 + tt;
        const bool kept = k < HOUT;
        const int kk = kept ? k : 0;
        const cpx<TW> zk = lds[kk * T + tt], zm = lds[(kk ? L - kk : 0) * T + tt];
        const TW h = k ? s : TW(0.5) * s;
        if (line < lines) {
            oa[line * hp + k] = kept ? cpx<TW>{(zk.re + zm.re) * h, (zk.im - zm.im) * h} : zero;
            ob[line * hp + k] = kept ? cpx<TW>{(zk.im + zm.im) * h, (zm.re - zk.re) * h} : zero;
        }
    }
    __syncthreads();
}

// VJP z pass.  in: fields 0..5 = u, w (padded half spectra, as the forward z
// pass reads them) and 6..8 = g, the cotangent of the truncated R2C output.
// out: 0..2 = GU = C2R^T (w x g), 3..5 = GW = C2R^T (g x u), truncated.
template <int L, typename TW>
__global__ __launch_bounds__((ZNLGeom<L, TW>::NT)) __attribute__((amdgpu_waves_per_eu(Waves<K_ZVJP, TW>::value))) void k_z_vjp(const cpx<TW> *__restrict__ in,
                                                                cpx<TW> *__restrict__ out, long long lines, int hp,
                                                                long long fs_in, long long fs_out, TW s_u, TW s_o,
                                                                const cpx<TW> *__restrict__ tw) {
    using P = ZNLP<L, TW>;
    using PL = typename P::radices;
    using RL = typename Reverse<PL>::type;
    constexpr int E = P::E;
    constexpr int T = ZNLGeom<L, TW>::T, NT = ZNLGeom<L, TW>::NT;
    constexpr int HIN = L / 3 + 1, HOUT = HIN;
    constexpr int TWN = tw_entries<L, PL>();
    __shared__ cpx<TW> lds[L * T];
    __shared__ cpx<TW> ltw[TWN];
    const int tid = threadIdx.x;
    const long long line0 = (long long)blockIdx.x * T;
    TwStage<TWN, NT, TW> tws;
    twiddles_load(tws, tw, tid);
    const int t = tid % T;
    const bool ok = line0 + t < lines;
    // lines past the end load the last line (valid memory); their results are not stored
    const long long base = (ok ? line0 + t : lines - 1) * hp;
    const TW one = TW(1), two = TW(2), half = TW(0.5);
    cpx<TW> v[E];
    TW g0[E], g1[E], r0[E], r1[E], r2[E], r3[E];

    // inverse pair transform of fields (fa, fb) (fb < 0: none) with per-half
    // DC factors: the z transpose of an R2C counts its k=0 term once where the
    // C2R extension would count it twice
    // (first: the LDS twiddle writes follow this transform's loads, see twiddles_load)
    auto inverse = [&](int fa, int fb, TW dca, TW dcb, bool first = false) {
        const gptr<cpx<TW>> pa = opaque(in) + fa * fs_in;
        ZRaw<TW> raw[E];
        if (fb >= 0) z_pair_issue<L, HIN, PL::first, T, NT, true>(raw, pa, opaque(in) + fb * fs_in, base, tid);
        else z_pair_issue<L, HIN, PL::first, T, NT, false>(raw, pa, pa, base, tid);
        if (first) twiddles_store(ltw, tws, tid);
        if (fb >= 0) z_pair_combine<L, HIN, PL::first, T, NT, true>(v, raw, tid);
        else z_pair_combine<L, HIN, PL::first, T, NT, false>(v, raw, tid);
        // position 0 is element 0 of thread column 0's first row
        if (ci_pos_in<L, PL::first, T, NT>(tid, 0) == 0) v[0] = {v[0].re * dca, v[0].im * dcb};
        zfft<L, +1, T, NT, PL>(v, lds, ltw, tid);
    };
    // Cotangent fields are paired only with cotangent fields: a complex FFT
    // of (a + i b) leaks roundoff of the larger of a, b into the other, and u, w
    // are many orders above g.
    // g = R2C^T gin: C2R with the DC counted once, times 1/2
    inverse(6, 7, two, two, true);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        g0[e] = v[e].re * half;
        g1[e] = v[e].im * half;
    }
    __syncthreads();
    inverse(8, -1, two, one);  // (gz, 0)
#pragma unroll
    for (int e = 0; e < E; ++e) r0[e] = v[e].re * half;  // g2
    __syncthreads();
    // W before U, holding W1, W2 and then U0 rather than partial products:
    // at most g (3 fields), two held fields and the transform pair are live
    inverse(4, 5, one, one);  // (wy, wz), held
#pragma unroll
    for (int e = 0; e < E; ++e) {
        r1[e] = v[e].re * s_u;  // W1
        r2[e] = v[e].im * s_u;  // W2
    }
    __syncthreads();
    inverse(3, 0, one, one);  // (wx, ux): gU complete; U0 held
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const TW w0 = v[e].re * s_u, u0 = v[e].im * s_u, gz = r0[e], w1 = r1[e], w2 = r2[e];
        const TW gu0 = w1 * gz - w2 * g1[e];     // gU0 = w1 g2 - w2 g1
        const TW gu1 = w2 * g0[e] - w0 * gz;     // gU1 = w2 g0 - w0 g2
        r1[e] = w0 * g1[e] - w1 * g0[e];         // gU2 = w0 g1 - w1 g0
        r3[e] = u0;
        v[e] = {gu0, gu1};
    }
    __syncthreads();
    zfft<L, -1, T, NT, RL>(v, lds, ltw, tid);
    __syncthreads();
    z_split_store<L, T, NT, RL, E, HOUT>(v, lds, tid, line0, lines, hp, out + 0 * fs_out, out + 1 * fs_out, s_o);
    // live: g0, g1, g2 = r0, gU2 = r1, U0 = r3
    inverse(1, 2, one, one);  // (uy, uz): gW complete
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const TW u1 = v[e].re * s_u, u2 = v[e].im * s_u, u0 = r3[e], gz = r0[e];
        const TW gw0 = g1[e] * u2 - gz * u1;  // gW0 = g1 U2 - g2 U1
        const TW gw1 = gz * u0 - g0[e] * u2;  // gW1 = g2 U0 - g0 U2
        r3[e] = g0[e] * u1 - g1[e] * u0;      // gW2 = g0 U1 - g1 U0
        v[e] = {gw0, gw1};
    }
    __syncthreads();
    zfft<L, -1, T, NT, RL>(v, lds, ltw, tid);
    __syncthreads();
    z_split_store<L, T, NT, RL, E, HOUT>(v, lds, tid, line0, lines, hp, out + 3 * fs_out, out + 4 * fs_out, s_o);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = {r1[e], r3[e]};
    zfft<L, -1, T, NT, RL>(v, lds, ltw, tid);
    __syncthreads();
    z_split_store<L, T, NT, RL, E, HOUT>(v, lds, tid, line0, lines, hp, out + 2 * fs_out, out + 5 * fs_out, s_o);
}

// The VJP z pass at L = 384 with one line per wavefront (the layouts of
// k_z_nl_w): the same fields, transforms and output as k_z_vjp.
template <typename TW, bool S64, int WPE, bool TC = false, bool PF = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_z_vjp_w(
    const cpx<TW> *__restrict__ in, cpx<TW> *__restrict__ out, long long lines, int hp, long long fs_in,
    long long fs_out, TW s_u, TW s_o, const cpx<TW> *__restrict__ tw) {
    using SP = ZwSpec<S64>;
    constexpr int NS = SP::NS, NP = S64 ? 8 : 6;  // spectral and physical slots
    constexpr int HOUT = zw::HIN, L = zw::L;
    __shared__ cpx<TW> lds[zw::NLDS];
    const int lane = threadIdx.x;
    const int j = lane < 48 ? lane : lane - 48;
    const bool wr = S64 || lane < 48;  // lanes that own spectral slots
    const long long base = (long long)blockIdx.x * hp;  // one line per block, grid = lines
    const gptr<cpx<TW>> gtw = (gptr<cpx<TW>>)tw;
    const ZwTw<TW, TC> tws(ldg(gtw, lane), ldg(gtw, 6 * (j & 7)));
    const cpx<TW> zero = {TW(0), TW(0)};
    const TW one = TW(1), two = TW(2), half = TW(0.5);
    cpx<TW> vs[NS], vp[NP];
    TW g0[NP], g1[NP], r0[NP], r1[NP], r2[NP], r3[NP];
    // The input pairs in order: (6, 7), (8, none), (4, 5), (3, 0), (1, 2).
    // With PF each next pair's loads (na, nb) are issued once the current pair
    // is combined, so they are in flight during its transform (and the
    // forward transforms between the inverse ones).
    ZRaw<TW> raw[NS];
    auto issue = [&](int fa, int fb) {
        const gptr<cpx<TW>> pa = opaque(in) + fa * fs_in;
        if (fb >= 0) zw_issue<S64, TW, true>(raw, pa, opaque(in) + fb * fs_in, base, lane, j);
        else zw_issue<S64, TW, false>(raw, pa, pa, base, lane, j);
    };
    if (PF) issue(6, 7);
    // inverse pair transform of fields (fa, fb) (fb < 0: none) with per-half
    // DC factors: the z transpose of an R2C counts its k=0 term once where the
    // C2R extension would count it twice
    auto inverse = [&](int fa, int fb, TW dca, TW dcb, int na = -1, int nb = -1) {
        if (!PF) issue(fa, fb);
        if (fb >= 0) zw_combine<S64, TW, true>(vs, raw, lane, j);
        else zw_combine<S64, TW, false>(vs, raw, lane, j);
        if (PF && na >= 0) {
            zw_after(vs);
            issue(na, nb);
        }
        // position 0 is slot 0 of lane 0 in both layouts
        if (lane == 0) vs[0] = {vs[0].re * dca, vs[0].im * dcb};
        if constexpr (S64) zw_64to48<+1>(vs, vp, lds, lane, j, tws);
        else zw_48to64<+1>(vs, vp, lds, lane, j, tws);
    };
    // forward DFT of vp split into the truncated half spectra of its real and
    // imaginary parts, times c_k s (c_0 = 1, c_k = 2): the transpose of the
    // unnormalised C2R.  Lines are written whole, entries [HOUT, hp) as zeros.
    constexpr int NK = 3, PLO = S64 ? 256 : 240;
    auto forward_split = [&](cpx<TW> *oa, cpx<TW> *ob) {
        if constexpr (S64) zw_48to64<-1>(vp, vs, lds, lane, j, tws);
        else zw_64to48<-1>(vp, vs, lds, lane, j, tws);
        __syncthreads();
        if (wr) {
#pragma unroll
            for (int s = NS - (S64 ? 2 : 3); s < NS; ++s) lds[SP::pos(lane, j, s) - PLO] = vs[s];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NK; ++s) {
            const int k = SP::pos(lane, j, s);
            const int li = (k >= 1 && k <= L - PLO) ? L - k - PLO : 0;
            const cpx<TW> zk = vs[s], zl = lds[li];
            const cpx<TW> zm = k ? zl : zk;
            const TW h = k ? s_o : half * s_o;
            const bool kept = k < HOUT;
            const cpx<TW> va = kept ? cpx<TW>{(zk.re + zm.re) * h, (zk.im - zm.im) * h} : zero;
            const cpx<TW> vb = kept ? cpx<TW>{(zk.im + zm.im) * h, (zm.re - zk.re) * h} : zero;
            if (wr && (s < NK - 1 || k < hp)) {
                oa[base + k] = va;
                ob[base + k] = vb;
            }
        }
        __syncthreads();
    };
    // Cotangent fields are paired only with cotangent fields: a complex FFT
    // of (a + i b) leaks roundoff of the larger of a, b into the other, and u, w
    // are many orders above g.
    // g = R2C^T gin: C2R with the DC counted once, times 1/2
    inverse(6, 7, two, two, 8, -1);
#pragma unroll
    for (int e = 0; e < NP; ++e) {
        g0[e] = vp[e].re * half;
        g1[e] = vp[e].im * half;
    }
    __syncthreads();
    inverse(8, -1, two, one, 4, 5);  // (gz, 0)
#pragma unroll
    for (int e = 0; e < NP; ++e) r0[e] = vp[e].re * half;  // g2
    __syncthreads();
    // W before U, holding W1, W2 and then U0 rather than partial products
    inverse(4, 5, one, one, 3, 0);  // (wy, wz), held
#pragma unroll
    for (int e = 0; e < NP; ++e) {
        r1[e] = vp[e].re * s_u;  // W1
        r2[e] = vp[e].im * s_u;  // W2
    }
    __syncthreads();
    inverse(3, 0, one, one, 1, 2);  // (wx, ux): gU complete; U0 held
#pragma unroll
    for (int e = 0; e < NP; ++e) {
        const TW w0 = vp[e].re * s_u, u0 = vp[e].im * s_u, gz = r0[e], w1 = r1[e], w2 = r2[e];
        const TW gu0 = w1 * gz - w2 * g1[e];  // gU0 = w1 g2 - w2 g1
        const TW gu1 = w2 * g0[e] - w0 * gz;  // gU1 = w2 g0 - w0 g2
        r1[e] = w0 * g1[e] - w1 * g0[e];      // gU2 = w0 g1 - w1 g0
        r3[e] = u0;
        vp[e] = {gu0, gu1};
    }
    __syncthreads();
    forward_split(out + 0 * fs_out, out + 1 * fs_out);
    // live: g0, g1, g2 = r0, gU2 = r1, U0 = r3
    inverse(1, 2, one, one);  // (uy, uz): gW complete
#pragma unroll
    for (int e = 0; e < NP; ++e) {
        const TW u1 = vp[e].re * s_u, u2 = vp[e].im * s_u, u0 = r3[e], gz = r0[e];
        const TW gw0 = g1[e] * u2 - gz * u1;  // gW0 = g1 U2 - g2 U1
        const TW gw1 = gz * u0 - g0[e] * u2;  // gW1 = g2 U0 - g0 U2
        r3[e] = g0[e] * u1 - g1[e] * u0;      // gW2 = g0 U1 - g1 U0
        vp[e] = {gw0, gw1};
    }
    __syncthreads();
    forward_split(out + 3 * fs_out, out + 4 * fs_out);
#pragma unroll
    for (int e = 0; e < NP; ++e) vp[e] = {r1[e], r3[e]};
    forward_split(out + 2 * fs_out, out + 5 * fs_out);
}

// C2R of three unpadded fields and the maximum of |v|^2 (CFL speed or |w|).
// The tile of the z pass: one-wave blocks of 2 lines (more for short lines),
// so each load instruction covers 32 or more consecutive positions of each
// line (tiles of 128-B rows across T lines read 32-B pieces of 16 lines per
// instruction in FP32).  Each block's maximum goes to one of kMaxSlots slots
// (block index modulo the count, one 128-B line each), and one small block
// reduces the slots: one atomic per block on a single address serialises.
// Tiles of 4 or 8 lines (2- and 4-wave blocks) made the CFL pass slower at
// 256^3 (mixed 0.255 -> 0.285 ms per step, FP64 0.467 -> 0.480).
template <int L, typename TW>
using ZMaxGeom = Geom<L, PlanSmall<L>::type::E, TW, 2 * 2 * (int)sizeof(TW)>;
// The standalone C2R and R2C along the contiguous axis use the z pass's tile:
// blocks of one wave holding T = 2 transform lines (more for short lines), the
// smaller-E plan and twiddles in LDS, so each load or store instruction covers
// 32 (or more) consecutive positions of each line (384^3: C2R 0.661 -> 0.567
// ms, R2C 0.762 -> 0.707 ms for the whole 3D transform).
template <int L, typename TW>
using ZLineGeom = Geom<L, PlanSmall<L>::type::E, TW, 2 * 2 * (int)sizeof(TW)>;

template <int L, typename TW>
__global__ __launch_bounds__((ZMaxGeom<L, TW>::NT)) void k_z_max(const cpx<TW> *__restrict__ in, long long lines,
                                                                 int hp, long long fs, TW s_in,
                                                                 const cpx<TW> *__restrict__ tw,
                                                                 unsigned long long *__restrict__ slots) {
    using P = typename PlanSmall<L>::type;
    using PL = typename P::radices;
    constexpr int E = P::E;
    constexpr int T = ZMaxGeom<L, TW>::T, NT = ZMaxGeom<L, TW>::NT;
    constexpr int HIN = L / 2 + 1;
    constexpr int TWN = tw_entries<L, PL>();
    __shared__ cpx<TW> lds[L * T];
    __shared__ cpx<TW> ltw[TWN];
    __shared__ double wmax[NT / 64];
    const int tid = threadIdx.x;
    TwStage<TWN, NT, TW> tws;
    twiddles_load(tws, tw, tid);
    const long long line0 = (long long)blockIdx.x * T;
    const int t = tid % T;
    const bool ok = line0 + t < lines;
    // lines past the end load the last line (valid memory) and are left out of the maximum
    const long long base = (ok ? line0 + t : lines - 1) * hp;
    cpx<TW> v[E];
    TW r[3][E];
    // the uz line's loads are in flight during the (ux, uy) pair's transform
    ZRaw<TW> raw[E];
    const gptr<cpx<TW>> f2 = opaque(in) + 2 * fs;
    z_pair_issue<L, HIN, PL::first, T, NT, true>(raw, opaque(in), opaque(in) + fs, base, tid);
    // the LDS twiddle writes follow the first pair's loads (see twiddles_load)
    twiddles_store(ltw, tws, tid);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (q == 0) {
            z_pair_combine<L, HIN, PL::first, T, NT, true>(v, raw, tid);
            z_pair_issue<L, HIN, PL::first, T, NT, false>(raw, f2, f2, base, tid);
        } else {
            z_pair_combine<L, HIN, PL::first, T, NT, false>(v, raw, tid);
        }
        fft_reg_lt<L, +1, T, NT, T, 1, PL>(v, lds, ltw, tid);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            r[2 * q][e] = v[e].re * s_in;
            if (q == 0) r[1][e] = v[e].im * s_in;
        }
        __syncthreads();
    }
    double m = 0.0;
    if (ok) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double a = r[0][e], b = r[1][e], c = r[2][e];
            m = fmax(m, a * a + b * b + c * c);
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((tid & 63) == 0) wmax[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < NT / 64; ++w) m = fmax(m, wmax[w]);
        atomicMax(&slots[(blockIdx.x % kMaxSlots) * kSlotStride], (unsigned long long)__double_as_longlong(m));
    }
}

// standalone C2R / R2C of one field along the contiguous axis: data lines
// (2s, 2s+1) share transform line s; a block holds T transform lines.
template <int L, typename TW>
using ZSGeom = ZLineGeom<L, TW>;

template <int L, typename TW>
__global__ __launch_bounds__((ZSGeom<L, TW>::NT)) void k_z_c2r(const cpx<TW> *__restrict__ in,
                                                               double *__restrict__ out, long long lines, int hp,
                                                               TW scale, const cpx<TW> *__restrict__ tw) {
    using P = typename PlanSmall<L>::type;
    using PL = typename P::radices;
    constexpr int E = P::E;
    constexpr int T = ZSGeom<L, TW>::T, NT = ZSGeom<L, TW>::NT;
    constexpr int HIN = L / 2 + 1;
    constexpr int TWN = tw_entries<L, PL>();
    __shared__ cpx<TW> lds[L * T];
    __shared__ cpx<TW> ltw[TWN];
    const int tid = threadIdx.x;
    TwStage<TWN, NT, TW> tws;
    twiddles_load(tws, tw, tid);
    const long long s = (long long)blockIdx.x * T + tid % T;
    const long long la = 2 * s, lb = la + 1;
    cpx<TW> v[E];
    // lines past the end load the last line (valid memory); b is zeroed past the end
    const gptr<cpx<TW>> src = opaque(in);
    {
        ZRaw<TW> ra[E], rb[E];
        z_pair_issue<L, HIN, PL::first, T, NT, false>(ra, src, src, (la < lines ? la : lines - 1) * hp, tid);
        z_pair_issue<L, HIN, PL::first, T, NT, false>(rb, src, src, (lb < lines ? lb : lines - 1) * hp, tid);
        // the LDS twiddle writes follow the lines' loads (see twiddles_load)
        twiddles_store(ltw, tws, tid);
        cpx<TW> vb[E];
        z_pair_combine<L, HIN, PL::first, T, NT, false>(v, ra, tid);
        z_pair_combine<L, HIN, PL::first, T, NT, false>(vb, rb, tid);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const cpx<TW> zb = lb < lines ? vb[e] : cpx<TW>{TW(0), TW(0)};
            v[e] = {v[e].re - zb.im, v[e].im + zb.re};  // X_a + i X_b
        }
    }
    fft_reg_lt<L, +1, T, NT, T, 1, PL>(v, lds, ltw, tid);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int p = ci_pos_out<L, PL::last, T, NT>(tid, e);
        if (la < lines) out[la * L + p] = double(v[e].re * scale);
        if (lb < lines) out[lb * L + p] = double(v[e].im * scale);
    }
}

template <int L, typename TW>
__global__ __launch_bounds__((ZSGeom<L, TW>::NT)) void k_z_r2c(const double *__restrict__ in,
                                                               cpx<TW> *__restrict__ out, long long lines, int hp,
                                                               const cpx<TW> *__restrict__ tw) {
    using P = typename PlanSmall<L>::type;
    using PL = typename P::radices;
    constexpr int E = P::E;
    constexpr int T = ZSGeom<L, TW>::T, NT = ZSGeom<L, TW>::NT;
    constexpr int HOUT = L / 2 + 1;
    constexpr int TWN = tw_entries<L, PL>();
    __shared__ cpx<TW> lds[L * T];
    __shared__ cpx<TW> ltw[TWN];
    const int tid = threadIdx.x;
    TwStage<TWN, NT, TW> tws;
    twiddles_load(tws, tw, tid);
    const long long pair0 = (long long)blockIdx.x * T;
    const int t = tid % T;
    const long long la = 2 * (pair0 + t), lb = la + 1;
    // lines past the end load the last line (valid memory) and are zeroed
    const long long ca = (la < lines ? la : lines - 1) * L, cb = (lb < lines ? lb : lines - 1) * L;
    const gptr<double> src = (gptr<double>)in;
    cpx<TW> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int p = ci_pos_in<L, PL::first, T, NT>(tid, e);
        const double a = src[ca + p], b = src[cb + p];
        v[e] = {la < lines ? TW(a) : TW(0), lb < lines ? TW(b) : TW(0)};
    }
    // the LDS twiddle writes follow the lines' loads (see twiddles_load)
    twiddles_store(ltw, tws, tid);
    fft_reg_lt<L, -1, T, NT, T, 1, PL>(v, lds, ltw, tid);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) lds[ci_pos_out<L, PL::last, T, NT>(tid, e) * T + t] = v[e];
    __syncthreads();
    const TW h = TW(0.5);
    const cpx<TW> zero = {TW(0), TW(0)};
    // lines are written whole, entries [HOUT, hp) as zeros
    for (int e = tid; e < T * hp; e += NT) {
        const int tt = e % T, k = e / T;
        const long long a = 2 * (pair0 + tt), b = a + 1;
        const bool kept = k < HOUT;
        const int kk = kept ? k : 0;
        const cpx<TW> zk = lds[kk * T + tt], zm = lds[(kk ? L - kk : 0) * T + tt];
        if (a < lines) out[a * hp + k] = kept ? cpx<TW>{(zk.re + zm.re) * h, (zk.im - zm.im) * h} : zero;
        if (b < lines) out[b * hp + k] = kept ? cpx<TW>{(zk.im + zm.im) * h, (zm.re - zk.re) * h} : zero;
    }
}

// ===========================================================================
// per-length launchers
// ===========================================================================
template <typename T>
inline dim3 blocks_for(long long n, int per) {
    return dim3((unsigned)((n + per - 1) / per));
}

template <int L, int LIN, int MODE, typename TI, typename TW>
int run_inv(const void *in, void *out, const void *tw, const Strided &g, int n, double ks, double scale,
            hipStream_t st) {
    using C = InvGeom<L, MODE, TW>;
    if (pad_stores<TW, C::T>() && g.ncols % C::T != 0) return 2;
    const unsigned nb = (unsigned)((g.ncols + C::T - 1) / C::T);
    hipLaunchKernelGGL((k_inv<L, LIN, MODE, TI, TW>), dim3(nb), dim3(C::NT), 0, st,
                       reinterpret_cast<const cpx<TI> *>(in), reinterpret_cast<cpx<TW> *>(out),
                       reinterpret_cast<const cpx<TW> *>(tw), g, n, ks, TW(scale));
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

template <int L, int LIN, typename TI, typename TW>
int run_inv_modes(int mode, const void *in, void *out, const void *tw, const Strided &g, int n, double ks,
                  double scale, hipStream_t st) {
    switch (mode) {
        case INV_PLAIN: return run_inv<L, LIN, INV_PLAIN, TI, TW>(in, out, tw, g, n, ks, scale, st);
        case INV_U3: return run_inv<L, LIN, INV_U3, TI, TW>(in, out, tw, g, n, ks, scale, st);
        case INV_U3DX: return run_inv<L, LIN, INV_U3DX, TI, TW>(in, out, tw, g, n, ks, scale, st);
        case INV_OMEGA: return run_inv<L, LIN, INV_OMEGA, TI, TW>(in, out, tw, g, n, ks, scale, st);
        case INV_UOMEGA: return run_inv<L, LIN, INV_UOMEGA, TI, TW>(in, out, tw, g, n, ks, scale, st);
    }
    return 2;
}

template <int L>
int launch_inv_L(int mode, int lin, Prec ip, Prec wp, const void *in, void *out, const void *tw,
                 const Strided &g, int n, double ks, double scale, hipStream_t st) {
    if (lin == L) {
        if (ip == F64 && wp == F64) return run_inv_modes<L, L, double, double>(mode, in, out, tw, g, n, ks, scale, st);
        if (ip == F64 && wp == F32) return run_inv_modes<L, L, double, float>(mode, in, out, tw, g, n, ks, scale, st);
        if (ip == F32 && wp == F32) return run_inv_modes<L, L, float, float>(mode, in, out, tw, g, n, ks, scale, st);
        return 2;
    }
    if constexpr (L % 3 == 0 && (2 * L / 3) % 2 == 0) {
        constexpr int LIN = 2 * L / 3;
        if (lin != LIN) return 2;
        if (mode != INV_U3DX && mode != INV_UOMEGA && mode != INV_U3) return 2;
        if (ip == F64 && wp == F64) {
            if (mode == INV_U3) return run_inv<L, LIN, INV_U3, double, double>(in, out, tw, g, n, ks, scale, st);
            return mode == INV_U3DX ? run_inv<L, LIN, INV_U3DX, double, double>(in, out, tw, g, n, ks, scale, st)
                                    : run_inv<L, LIN, INV_UOMEGA, double, double>(in, out, tw, g, n, ks, scale, st);
        }
        if (ip == F64 && wp == F32 && mode == INV_U3DX)
            return run_inv<L, LIN, INV_U3DX, double, float>(in, out, tw, g, n, ks, scale, st);
        if (ip == F64 && wp == F32 && mode == INV_U3)
            return run_inv<L, LIN, INV_U3, double, float>(in, out, tw, g, n, ks, scale, st);
        if (ip == F32 && wp == F32 && mode == INV_UOMEGA)
            return run_inv<L, LIN, INV_UOMEGA, float, float>(in, out, tw, g, n, ks, scale, st);
        if (ip == F32 && wp == F32 && mode == INV_U3DX)
            return run_inv<L, LIN, INV_U3DX, float, float>(in, out, tw, g, n, ks, scale, st);
        if (ip == F32 && wp == F32 && mode == INV_U3)
            return run_inv<L, LIN, INV_U3, float, float>(in, out, tw, g, n, ks, scale, st);
    }
    return 2;
}

template <int L, int LOUT, int MODE, int NF, int STG, typename TW>
int run_fwd(const void *in, void *out, const void *tw, const Strided &g, int n, double ks, double scale,
            const StageArgs *sa, hipStream_t st) {
    using C = FwdGeom<L, MODE, TW>;
    if ((MODE == FWD_PLAIN || MODE == FWD_VJPY) && NF != 1 && pad_stores<TW, C::T>() && g.ncols % C::T != 0) return 2;
    const unsigned nb = (unsigned)((g.ncols + C::T - 1) / C::T);
    StageArgs a{};
    if (sa) a = *sa;
    hipLaunchKernelGGL((k_fwd<L, LOUT, MODE, NF, STG, TW>), dim3(nb), dim3(C::NT), 0, st,
                       reinterpret_cast<const cpx<TW> *>(in), reinterpret_cast<cpx<TW> *>(out),
                       reinterpret_cast<const cpx<TW> *>(tw), g, n, ks, TW(scale), a);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

// stage passes: the runtime stage selects the kernel instance
template <int L, int LOUT, typename TW>
int run_fwd_stage(const void *in, void *out, const void *tw, const Strided &g, int n, double ks, double scale,
                  const StageArgs *sa, hipStream_t st) {
    switch (sa->stage) {
        case ST_RAW: return run_fwd<L, LOUT, FWD_STAGE, 3, ST_RAW, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case ST_1: return run_fwd<L, LOUT, FWD_STAGE, 3, ST_1, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case ST_2: return run_fwd<L, LOUT, FWD_STAGE, 3, ST_2, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case ST_3: return run_fwd<L, LOUT, FWD_STAGE, 3, ST_3, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case ST_4: return run_fwd<L, LOUT, FWD_STAGE, 3, ST_4, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case RK_1: return run_fwd<L, LOUT, FWD_STAGE, 3, RK_1, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case RK_2: return run_fwd<L, LOUT, FWD_STAGE, 3, RK_2, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case RK_3: return run_fwd<L, LOUT, FWD_STAGE, 3, RK_3, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case RK_4: return run_fwd<L, LOUT, FWD_STAGE, 3, RK_4, TW>(in, out, tw, g, n, ks, scale, sa, st);
    }
    return 2;
}
template <int L, int LOUT, typename TW>
int run_fwd_vjp(const void *in, void *out, const void *tw, const Strided &g, int n, double ks, double scale,
                const StageArgs *sa, hipStream_t st) {
    switch (sa->stage) {
        case AS_RAW: return run_fwd<L, LOUT, FWD_VJP, 5, AS_RAW, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case AS_1: return run_fwd<L, LOUT, FWD_VJP, 5, AS_1, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case AS_2: return run_fwd<L, LOUT, FWD_VJP, 5, AS_2, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case AS_3: return run_fwd<L, LOUT, FWD_VJP, 5, AS_3, TW>(in, out, tw, g, n, ks, scale, sa, st);
        case AS_4: return run_fwd<L, LOUT, FWD_VJP, 5, AS_4, TW>(in, out, tw, g, n, ks, scale, sa, st);
    }
    return 2;
}

template <int L>
int launch_fwd_L(int mode, int lout, int nf, Prec wp, const void *in, void *out, const void *tw,
                 const Strided &g, int n, double ks, double scale, const StageArgs *sa, hipStream_t st) {
    if (lout == L) {
        if (mode == FWD_PLAIN && nf == 1 && wp == F64)
            return run_fwd<L, L, FWD_PLAIN, 1, 0, double>(in, out, tw, g, n, ks, scale, sa, st);
        return 2;
    }
    if constexpr (L % 3 == 0 && (2 * L / 3) % 2 == 0) {
        constexpr int LOUT = 2 * L / 3;
        if (lout != LOUT) return 2;
        if (mode == FWD_PLAIN && nf == 3) {
            if (wp == F64) return run_fwd<L, LOUT, FWD_PLAIN, 3, 0, double>(in, out, tw, g, n, ks, scale, sa, st);
            return run_fwd<L, LOUT, FWD_PLAIN, 3, 0, float>(in, out, tw, g, n, ks, scale, sa, st);
        }
        if (mode == FWD_STAGE && nf == 3 && sa) {
            if (wp == F64) return run_fwd_stage<L, LOUT, double>(in, out, tw, g, n, ks, scale, sa, st);
            return run_fwd_stage<L, LOUT, float>(in, out, tw, g, n, ks, scale, sa, st);
        }
        if (mode == FWD_PLAIN && nf == 6) {
            if (wp == F64) return run_fwd<L, LOUT, FWD_PLAIN, 6, 0, double>(in, out, tw, g, n, ks, scale, sa, st);
            return run_fwd<L, LOUT, FWD_PLAIN, 6, 0, float>(in, out, tw, g, n, ks, scale, sa, st);
        }
        if (mode == FWD_VJPY && nf == 6) {
            if (wp == F64) return run_fwd<L, LOUT, FWD_VJPY, 6, 0, double>(in, out, tw, g, n, ks, scale, sa, st);
            return run_fwd<L, LOUT, FWD_VJPY, 6, 0, float>(in, out, tw, g, n, ks, scale, sa, st);
        }
        if (mode == FWD_VJP && nf == 5 && sa) {
            if (wp == F64) return run_fwd_vjp<L, LOUT, double>(in, out, tw, g, n, ks, scale, sa, st);
            return run_fwd_vjp<L, LOUT, float>(in, out, tw, g, n, ks, scale, sa, st);
        }
    }
    return 2;
}

template <int L>
int launch_z_nl_L(Prec wp, const void *in, void *out, long long lines, int hp, long long fs_in, long long fs_out,
                  double s_in, double s_out, const void *tw, hipStream_t st) {
    if constexpr (L == 384) {
        // one line per wave (k_z_nl_w); measured per launch at 256^3 against
        // k_z_nl (0.758 / 0.384 ms FP64 / FP32) on the same boxes:
        //   FP64: spectral side in layout 48, no prefetch, a 3-wave floor:
        //   0.659 ms with one line per block (162 VGPRs), 0.61-0.625 with two
        //   (c2 paired; 166 VGPRs).  Layout 64 spills 44 B at that floor
        //   (0.742 ms), prefetch at two waves ran 0.63-0.70, and issuing the
        //   second line's loads during the first line's forward transforms
        //   measured the same.  Cached twiddle powers (as FP32) spill 40 B
        //   at this floor.
        //   FP32: spectral side in layout 64 with prefetch, one line per block
        //   (92 VGPRs, five waves): 0.326 ms; two lines per block 0.331,
        //   without prefetch 0.347, layout 48 0.356.  The twiddle powers formed
        //   once per wave (99 VGPRs, four waves): 0.326 -> 0.321 ms; a 5-wave
        //   floor spills 24 B and ran 0.355 ms; two lines per block with
        //   their c2 paired (105 VGPRs, four waves) 0.327 ms.
        if (wp == F64) {
            if (lines % 2 == 0)
                hipLaunchKernelGGL((k_z_nl_w<double, false, false, 3, 2>), dim3((unsigned)(lines / 2)), dim3(64), 0,
                                   st, reinterpret_cast<const cpx<double> *>(in), reinterpret_cast<cpx<double> *>(out),
                                   lines, hp, fs_in, fs_out, s_in, s_out, reinterpret_cast<const cpx<double> *>(tw));
            else
                hipLaunchKernelGGL((k_z_nl_w<double, false, false, 3, 1>), dim3((unsigned)lines), dim3(64), 0, st,
                                   reinterpret_cast<const cpx<double> *>(in), reinterpret_cast<cpx<double> *>(out),
                                   lines, hp, fs_in, fs_out, s_in, s_out, reinterpret_cast<const cpx<double> *>(tw));
        } else {
            hipLaunchKernelGGL((k_z_nl_w<float, true, true, 4, 1, true>), dim3((unsigned)lines), dim3(64), 0, st,
                               reinterpret_cast<const cpx<float> *>(in), reinterpret_cast<cpx<float> *>(out), lines,
                               hp, fs_in, fs_out, float(s_in), float(s_out), reinterpret_cast<const cpx<float> *>(tw));
        }
        return hipGetLastError() == hipSuccess ? 0 : 3;
    }
    if constexpr (L % 3 == 0 && (2 * L / 3) % 2 == 0) {
        if (wp == F64) {
            using G = ZNLGeom<L, double>;
            hipLaunchKernelGGL((k_z_nl<L, double>), blocks_for<void>(lines, G::T), dim3(G::NT), 0, st,
                               reinterpret_cast<const cpx<double> *>(in), reinterpret_cast<cpx<double> *>(out), lines,
                               hp, fs_in, fs_out, s_in, s_out, reinterpret_cast<const cpx<double> *>(tw));
        } else {
            using G = ZNLGeom<L, float>;
            hipLaunchKernelGGL((k_z_nl<L, float>), blocks_for<void>(lines, G::T), dim3(G::NT), 0, st,
                               reinterpret_cast<const cpx<float> *>(in), reinterpret_cast<cpx<float> *>(out), lines,
                               hp, fs_in, fs_out, float(s_in), float(s_out), reinterpret_cast<const cpx<float> *>(tw));
        }
        return hipGetLastError() == hipSuccess ? 0 : 3;
    }
    return 2;
}

template <int L>
int launch_z_vjp_L(Prec wp, const void *in, void *out, long long lines, int hp, long long fs_in, long long fs_out,
                   double s_u, double s_o, const void *tw, hipStream_t st) {
    if constexpr (L == 384) {
        // one line per wave (k_z_vjp_w); measured at 256^3 by the adjoint's
        // reverse step against k_z_vjp on the same box: FP64 17.8 -> 17.2 ms
        // with the spectral side in layout 48 at two waves (196 VGPRs; a
        // 3-wave floor spills 156 B, 108 B with W1, W2 held in LDS, and ran
        // 18.7 ms); mixed 10.94 -> 10.09 ms with the spectral side in layout
        // 64 (112 VGPRs, four waves; layout 48 10.33 ms).  The twiddle powers
        // formed once per wave keep those occupancies (206 / 119 VGPRs):
        // reverse step 16.84 -> 16.71 ms FP64, 9.98 -> 9.91 ms mixed.  At two
        // waves the FP64 kernel has room to load each next field pair during
        // the current one's transforms (248 VGPRs, no spill): 16.91 -> 16.55
        // ms; in FP32 the same prefetch spills 32 B at four waves and ran
        // 9.91 -> 10.45 ms.
        if (wp == F64)
            hipLaunchKernelGGL((k_z_vjp_w<double, false, 2, true, true>), dim3((unsigned)lines), dim3(64), 0, st,
                               reinterpret_cast<const cpx<double> *>(in), reinterpret_cast<cpx<double> *>(out), lines,
                               hp, fs_in, fs_out, s_u, s_o, reinterpret_cast<const cpx<double> *>(tw));
        else
            hipLaunchKernelGGL((k_z_vjp_w<float, true, 4, true>), dim3((unsigned)lines), dim3(64), 0, st,
                               reinterpret_cast<const cpx<float> *>(in), reinterpret_cast<cpx<float> *>(out), lines,
                               hp, fs_in, fs_out, float(s_u), float(s_o), reinterpret_cast<const cpx<float> *>(tw));
        return hipGetLastError() == hipSuccess ? 0 : 3;
    }
    if constexpr (L % 3 == 0 && (2 * L / 3) % 2 == 0) {
        if (wp == F64) {
            using G = ZNLGeom<L, double>;
            hipLaunchKernelGGL((k_z_vjp<L, double>), blocks_for<void>(lines, G::T), dim3(G::NT), 0, st,
                               reinterpret_cast<const cpx<double> *>(in), reinterpret_cast<cpx<double> *>(out), lines,
                               hp, fs_in, fs_out, s_u, s_o, reinterpret_cast<const cpx<double> *>(tw));
        } else {
            using G = ZNLGeom<L, float>;
            hipLaunchKernelGGL((k_z_vjp<L, float>), blocks_for<void>(lines, G::T), dim3(G::NT), 0, st,
                               reinterpret_cast<const cpx<float> *>(in), reinterpret_cast<cpx<float> *>(out), lines,
                               hp, fs_in, fs_out, float(s_u), float(s_o), reinterpret_cast<const cpx<float> *>(tw));
        }
        return hipGetLastError() == hipSuccess ? 0 : 3;
    }
    return 2;
}

template <int L>
int launch_z_max_L(Prec wp, const void *in, long long lines, int hp, long long fs, double s_in, const void *tw,
                   unsigned long long *slots, hipStream_t st) {
    if (wp == F64) {
        using G = ZMaxGeom<L, double>;
        hipLaunchKernelGGL((k_z_max<L, double>), blocks_for<void>(lines, G::T), dim3(G::NT), 0, st,
                           reinterpret_cast<const cpx<double> *>(in), lines, hp, fs, s_in,
                           reinterpret_cast<const cpx<double> *>(tw), slots);
    } else {
        using G = ZMaxGeom<L, float>;
        hipLaunchKernelGGL((k_z_max<L, float>), blocks_for<void>(lines, G::T), dim3(G::NT), 0, st,
                           reinterpret_cast<const cpx<float> *>(in), lines, hp, fs, float(s_in),
                           reinterpret_cast<const cpx<float> *>(tw), slots);
    }
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

template <int L>
int launch_z_c2r_L(const void *in, double *out, long long lines, int hp, double scale, const void *tw,
                   hipStream_t st) {
    using G = ZSGeom<L, double>;
    hipLaunchKernelGGL((k_z_c2r<L, double>), blocks_for<void>((lines + 1) / 2, G::T), dim3(G::NT), 0, st,
                       reinterpret_cast<const cpx<double> *>(in), out, lines, hp, scale,
                       reinterpret_cast<const cpx<double> *>(tw));
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

template <int L>
int launch_z_r2c_L(const double *in, void *out, long long lines, int hp, const void *tw, hipStream_t st) {
    using G = ZSGeom<L, double>;
    hipLaunchKernelGGL((k_z_r2c<L, double>), blocks_for<void>((lines + 1) / 2, G::T), dim3(G::NT), 0, st, in,
                       reinterpret_cast<cpx<double> *>(out), lines, hp, reinterpret_cast<const cpx<double> *>(tw));
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

}  // namespace vk
