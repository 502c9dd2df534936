// sparc_colseg.hip -- segmented column pass of the 2D option-17 solver.
//
// Same work as the column half of one RK stage of step.f90:48-106 /
// physics2d.f90:943-1094: the z-derivatives of xi_z, v_z, v_x and p by dbyd2
// (dbyd2.f:8-200, mtridag.f:95-112), the z-PML, divergence, pressure and
// density perturbation, RHS v_z with the source and flow terms, the zeroed
// rows and columns, the RK stage update and, after the last stage,
// FILTER_VARS_2D (physics2d.f90:509-556).
//
// Every column is split into P <= kSegMax segments of exactly MS rows; one
// lane holds one segment of one field of one column, and a block covers TC
// columns (k_colsplit below).  Each z-derivative is: dbyd2 right-hand side
// from a 7-row window, local Thomas solve of the segment, exchange of the
// segment's end values through LDS, and the exact correction
// y - L_{s-1} V - F_{s+1} W.  Loads use clamped row indices and 32-bit byte
// offsets from wave-uniform bases; stores of padding rows are masked off.
#include <type_traits>

#include "sparc_device.h"
#include "sparc_launch.h"
#include "sparc_seg.h"

namespace sparc {

namespace {

enum ZRow { Z_STRETCH = 0, Z_G = 1, Z_RHO = 17, Z_GRHO = 18, Z_GP = 19, Z_RINV = 20 };
__device__ __forceinline__ int zrow(int slot, int k) { return 2 + 5 * slot + k; }  // k: 0 lowr 1 rbl 2 gl 3 V 4 W

}  // namespace

// z tables staged in LDS: the segment tables of the three dbyd2 table sets
// (ibc 1, 2, 3; rows 5 slot + k, k: lowr, rbl, gl, V, W) of the first, an
// interior and the last segment, [kZS][3][MS], then the six epilogue
// profiles [kZF][nzp] in this order (stage_prow).
enum PfRow { F_STR = 0, F_G, F_RHO, F_GRHO, F_GP, F_RINV };
static_assert(F_RINV + 1 == kZF, "staged z profiles");
// The block's z-PML data staged in LDS, [kPL][nzpml][TC]: az, bz, fz and the
// two auxiliary fields (psivz, psip), stage value and base.
enum PRow { P_AZ = 0, P_BZ, P_FZ, P_T0, P_B0, P_T1, P_B1, kPL };

// Dynamic LDS of one block, in this order: the staged region (the three
// banded interface inverses [3][2P][4hw+4], the segment tables [kZS][3][ms],
// the z profiles [kZF][nzp], dbyd2's kCB boundary-row coefficients, the z-PML data [kPL][nzpml][TC], the mode-3 source z profile [nzp] (zero on
// padding rows; only staged in that mode); filled in one pass at the start), the four
// end-value exchange buffers [4][TC][2P + 4hw + 2] (each column's 2P end
// values between 2hw and 2hw+2 zeros), and three [TC][nzp] row buffers: the
// group-1 and group-0 parts of RHS v_z and p.
inline size_t colsplit_staged(int tc, int P, int hw, int nzp, int nzpml, int ms)
{
    // rounded up to an even count, so that the end-value buffers after it
    // start on 16 bytes (the interface correction reads them in pairs)
    const size_t n = (size_t)3 * 2 * P * (4 * hw + 4) + (size_t)kZS * 3 * ms + (size_t)kZF * nzp + (size_t)kCB +
                     (size_t)kPL * nzpml * tc + (size_t)nzp;
    return (n + 1) & ~(size_t)1;
}
inline size_t colsplit_lds_bytes(int tc, int P, int hw, int nzp, int nzpml, int ms)
{
    return sizeof(double) *
           (colsplit_staged(tc, P, hw, nzp, nzpml, ms) + (size_t)4 * tc * yb_len(P, hw) + (size_t)3 * tc * nzp);
}

// ---------------------------------------------------------------------------
// The column pass, one lane per (column, segment, field): the three
// independent first derivatives (xi_z, v_z, v_x) run in three lane groups of
// G = roundup(TC*P, 64) <= 128 threads each (blockDim = 3G, so every wave
// holds one field), which splits the window loads, solves, corrections and epilogue
// work of the first half three ways.  Group 0 (xi_z) then does d p/dz and
// the v_z update; group 1 writes the xi_z update and passes its part of RHS
// v_z through LDS; group 2 writes RHS v_x and the xi_x update.  After the
// last stage each group filters one field.  Store/load order across groups:
// xi_z (TN[1]) is written only after the first barrier, when every lane's
// window reads of T[1] are complete (T and TN are one buffer after the first
// stage); v_z (TN[3]) after the third.
// Every table the pass reads comes from LDS, so no lane waits on global
// memory for them once the staging barrier has passed.

// Loads and stores at a 32-bit byte offset from a wave-uniform base address:
// the compiler then uses the SGPR-base + VGPR-offset addressing form, with no
// 64-bit address arithmetic per access.  The host checks that every array
// indexed this way spans less than 2^31 bytes.
#define SPARC_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ double ldb(const double *base, unsigned ob)
{
    return *(const SPARC_GLOBAL double *)((const SPARC_GLOBAL char *)base + ob);
}
__device__ __forceinline__ void stb(double *base, unsigned ob, double v)
{
    *(SPARC_GLOBAL double *)((SPARC_GLOBAL char *)base + ob) = v;
}
// A value every lane of the wave holds, moved to a scalar register.
template <typename T>
__device__ __forceinline__ T *wave_ptr(T *p)
{
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}

template <int MS, int LAST>
__global__ __launch_bounds__(384) void k_colsplit(Dims d, Bg bg, DbyCoef cf, SegTab sg, State st,
                                                  ColArgs ca, Src src, Clock clk, Sink sink, int TC)
{
    extern __shared__ double lds_dyn[];
    const int P = sg.P, n2P = 2 * P, nzp = sg.nzp;
    const int nband = n2P * (4 * sg.hw + 4);
    const int nzpml = d.nzpml;
    double *const mv = lds_dyn;                        // [3][nband] banded interface inverses
    double *const zl = mv + 3 * nband;                 // [kZS][3][MS] segment tables
    double *const zf = zl + kZS * 3 * MS;              // [kZF][nzp] z profiles
    double *const cbl = zf + kZF * nzp;                // [kCB] dbyd2 boundary-row coefficients
    double *const pl = cbl + kCB;                      // [kPL][nzpml][TC] z-PML data
    double *const zpl = pl + kPL * nzpml * TC;         // [nzp] mode-3 source z profile
    const int nstaged = (3 * nband + kZS * 3 * MS + kZF * nzp + kCB + kPL * nzpml * TC + nzp + 1) & ~1;
    const int LY = n2P + 4 * sg.hw + 2, hw2 = 2 * sg.hw;
    double *const yb0 = lds_dyn + nstaged;             // [4][TC][LY] end values: field groups 0-2, d p/dz
    double *const wl = yb0 + 4 * TC * LY;              // [TC][nzp] group-1 part of RHS v_z,
                                                       // then [TC][nzp] group 0's part, [TC][nzp] p
    const int tid = threadIdx.x, G = blockDim.x / 3;
    const int f = __builtin_amdgcn_readfirstlane(tid / G);   // the field group: one per wave
    // end-value buffers by offset from the LDS base (an array of pointers
    // indexed by f would turn these accesses into flat ones)
    double *const ybf = yb0 + f * TC * LY, *const yb3 = yb0 + 3 * TC * LY;
    const int l = tid - f * G;
    const int c = l % TC, s = l / TC;
    const bool act = s < P;
    const int sc = act ? s : 0;
    const int a = sc * MS;
    const bool first = act && s == 0;
    const bool lastseg = a + MS > d.nz - 2;            // holds rows nz-2 and nz-1
    int bx = blockIdx.x, by = blockIdx.y;
    {
        const int n = gridDim.x * gridDim.y;
        if ((n & 7) == 0) {
            const int id = blockIdx.x + gridDim.x * blockIdx.y;
            const int lg = (id & 7) * (n >> 3) + (id >> 3);
            bx = lg % gridDim.x;
            by = lg / gridDim.x;
        }
    }
    const int x = bx * TC + c, b = by;
    const int nx = d.nx, nz = d.nz;
    const int nxb = nx * 8;                            // bytes per row
    const int colb = (b * nz * nx + x) * 8;            // byte offset of (b, row 0, x) in an [S][nz][nx] array
    const int gcolb = x * 8;                           // ... of (row 0, x) in an [nz][nx] array
    const int lastb = (nz - 1) * nxb;
    const int time = clock_t_(clk);
    const int nb = d.npmlbot, ntop = d.npmltop;
    const int pz0b = (b * nzpml * nx + x) * 8;
    const int pz1b = ((d.S + b) * nzpml * nx + x) * 8;
    const double *zs = zf + a;                         // row a of the staged z profiles
    // this segment's tables: those of the first, an interior or the last
    // segment (every interior segment's are the same), row r at zt[r * 3 MS + i]
    const double *zt = zl + seg_kind(sc, P) * MS;
    auto plv = [&](int k, int pj) -> double { return pl[(k * nzpml + pj) * TC + c]; };

    // byte offset of row r of this column, r clamped to the column
    auto rowc = [&](int r) -> unsigned { return (unsigned)(min(max(r * nxb, 0), lastb)); };
    auto valid = [&](int i) -> bool { return act && a + i < nz; };
    auto pmlrow = [&](int j) -> int {
        const int jc = min(j, nz - 1);
        return jc < nb ? jc : min(nzpml - 1, nb + max(jc - (nz - ntop), 0));
    };
    auto is_pml = [&](int j) -> bool { return j < nb || (j >= nz - ntop && j < nz); };
    // The epilogues take their z-PML form for the whole wave when any lane's
    // segment meets a PML layer: that form leaves rows outside the layers as
    // the plain form does (the PML terms are selected off and the auxiliary
    // stores masked), and a per-lane choice would run both forms one after
    // the other in every wave that holds a layer.
    const bool segpml = __ballot(a < nb || a + MS > nz - ntop) != 0;

    // --- first derivatives: group 0 d xi_z/dz (ibc 3, top row preset),
    //     group 1 d v_z/dz (ibc 1), group 2 d v_x/dz (ibc 1)
    const double *F = wave_ptr(f == 0 ? ca.T[1] : (f == 1 ? ca.T[3] : ca.T[2]));
    double win[MS + 6];
#pragma unroll
    for (int k = 0; k < MS + 6; k++) win[k] = ldb(F, colb + rowc(a - 3 + k));
    const double topv = f == 0 ? ldb(st.topbc, (b * nx + x) * 8) : 0.0;
    const double r0src = f == 0 ? ldb(bg.rho0_src, gcolb) : 1.0;   // the source lanes' rho0 at the source row
    double e[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (lastseg) {
#pragma unroll
        for (int k = 0; k < 5; k++) e[k] = ldb(F, colb + (nz - 5 + k) * nxb);
    }
    {
        // zero the pads of the end-value buffers
        const int npad = 4 * TC, padw = LY - n2P;
        for (int q = tid; q < npad * padw; q += blockDim.x) {
            const int r = q / padw, u = q - r * padw;
            yb0[r * LY + (u < hw2 ? u : n2P + u)] = 0.0;
        }
    }
    {
        // the staged region in one pass for the usual shapes: the inverses
        // and z tables are one contiguous copy of sg.stage (kStage loads per
        // thread), the z-PML data come from their own arrays (two elements per
        // thread and kind); every load of a round is issued before its stores
        constexpr int kStage = 32;
        const int nA = 3 * nband + kZS * 3 * MS + kZF * nzp + kCB, npl = nzpml * TC, stride = blockDim.x;
        const int x0 = bx * TC;
        const int rounds = max((nA + kStage * stride - 1) / (kStage * stride),
                               (max(npl, nzp) + 2 * stride - 1) / (2 * stride));
        const bool zsrc = src.mode == 3;
        const double *const stg = wave_ptr(sg.stage);
        for (int r0 = 0; r0 < rounds; r0++) {
            double va[kStage], vb[kPL][2], vz[2] = {0.0, 0.0};
            const int qa0 = r0 * kStage * stride + tid, qb0 = r0 * 2 * stride + tid;
#pragma unroll
            for (int u = 0; u < kStage; u++) va[u] = ldb(stg, (unsigned)min(qa0 + u * stride, nA - 1) * 8u);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int q = min(qb0 + h * stride, npl - 1);
                const int pj = q / TC, xc = x0 + (q - pj * TC);
                const unsigned gq = (unsigned)(pj * nx + xc) * 8u;
                const unsigned s0 = (unsigned)((b * nzpml + pj) * nx + xc) * 8u;
                const unsigned s1 = (unsigned)(((d.S + b) * nzpml + pj) * nx + xc) * 8u;
                vb[P_AZ][h] = ldb(bg.az, gq);
                vb[P_BZ][h] = ldb(bg.bz, gq);
                vb[P_FZ][h] = ldb(bg.fz, gq);
                vb[P_T0][h] = ldb(st.pz_tmp, s0);
                vb[P_B0][h] = ldb(st.pz_base, s0);
                vb[P_T1][h] = ldb(st.pz_tmp, s1);
                vb[P_B1][h] = ldb(st.pz_base, s1);
            }
            if (zsrc) {
#pragma unroll
                for (int h = 0; h < 2; h++) vz[h] = ldb(src.zprof, (unsigned)min(qb0 + h * stride, nz - 1) * 8u);
            }
#pragma unroll
            for (int u = 0; u < kStage; u++)
                if (qa0 + u * stride < nA) lds_dyn[qa0 + u * stride] = va[u];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int q = qb0 + h * stride;
                if (q < npl) {
#pragma unroll
                    for (int k = 0; k < kPL; k++) pl[k * npl + q] = vb[k][h];
                }
                if (zsrc && q < nzp) zpl[q] = q < nz ? vz[h] : 0.0;
            }
        }
    }
    __syncthreads();
    // source value of this column at its injection row (physics2d.f90:1051-1061)
    double srcv = 0.0;
    const int isrc = src.row - a;
    if (f == 0 && act && isrc >= 0 && isrc < MS && time <= src.limit) {
        if (src.mode == 1) {
            // both loads issued before either is used (clamped time index)
            double fsr = src.fwd[min(max(time, 0), max(src.nfwd - 1, 0))];
            double vrx = src.vr[b * nx + x];
            asm volatile("" : "+v"(fsr), "+v"(vrx));
            const double fs = (time >= 0 && time < src.nfwd) ? fsr : 0.0;
            srcv = vrx * fs * src.delta_width / r0src;
        } else if (src.mode == 2) {
            srcv = lagrange_forcing(src, b, x, nx, time) * src.delta_width / r0src;
        }
    }
    const double pamp = (f == 0 && act) ? profile_amp(src, b, x, nx, a, MS, time) : 0.0;
    const int slot_f = f == 0 ? 2 : 0;                 // dbyd2 table set: ibc 3 or ibc 1
    double Y[MS], own[MS];
    const double *const cbt = cbl;                      // LDS: read only by the end segments' lanes
    seg_rhs<MS>([&](int k) { return win[k]; }, e, Y, a, nz, first, lastseg, f == 0 ? 3 : 1, 0.0, topv, cf.c1,
                cf.c2, cbt);
#pragma unroll
    for (int i = 0; i < MS; i++) own[i] = win[3 + i];
    const double *const zt_f = zt + 5 * slot_f * 3 * MS;   // this group's table set
    auto gtab = [&](int t, int i) { return zt_f[t * 3 * MS + i]; };
    seg_solve<MS>(Y, gtab);
    if (act) { ybf[c * LY + hw2 + 2 * s] = Y[0]; ybf[c * LY + hw2 + 2 * s + 1] = Y[MS - 1]; }
    // the epilogue's inputs that no lane of this kernel writes, loaded
    // before the barrier so that they arrive during the exchange and the
    // correction: group 0 d xi_x/dx, c^2 rho0 and the v_z base; groups 1, 2
    // their x-derivative, the flow and the base of their update
    double pa[MS], pb[MS], pc[MS];
    if (f == 0) {
#pragma unroll
        for (int i = 0; i < MS; i++) {
            const unsigned r = rowc(a + i);
            pa[i] = ldb(st.dxixdx, colb + r);
            pb[i] = ldb(bg.c2rho0, gcolb + r);
            pc[i] = ldb(st.a[3], colb + r);
        }
    }
    __syncthreads();
    if (act)
        seg_correct<MS>(Y, s, P, sg.hw, ybf + c * LY, mv + (f == 0 ? 2 * nband : 0), gtab);

    double *const w0 = wl + (TC + c) * nzp + a;       // group 0: -rho'/rho0 g + source
    double *const pcol = wl + (2 * TC + c) * nzp;     // p of this column (group 0)
    if (f == 0) {
        // z-PML (:957-967), divergence, p and rho' (:990-998, :1043)
        auto epi = [&](auto pml_c) {
            constexpr bool PML = decltype(pml_c)::value;
#pragma unroll
            for (int i = 0; i < MS; i++) {
                double lpsi = 0.0, laz = 0.0, lbz = 0.0, lpb = 0.0, lfz = 0.0;
                if (PML) {
                    const int pj = pmlrow(a + i);
                    lpsi = plv(P_T0, pj);
                    lpb = plv(P_B0, pj);
                    laz = plv(P_AZ, pj);
                    lbz = plv(P_BZ, pj);
                    lfz = plv(P_FZ, pj);
                }
                const int j = a + i;
                const double str = zs[F_STR * nzp + i], r0 = zs[F_RHO * nzp + i];
                const double rinv = zs[F_RINV * nzp + i], gz = zs[F_G * nzp + i];
                double dz = Y[i] * str;
                double fdec = 1.0;
                if (PML) {
                    const bool pm = is_pml(j);
                    fdec = pm ? lfz : 1.0;
                    const double prhs = laz * dz + lbz * lpsi;
                    dz = pm ? dz + lpsi : dz;
                    const double pnew = lpb + ca.bnext * prhs;
                    if (pm && act) {
                        const unsigned qp = pz0b + pmlrow(j) * nxb;
                        stb(st.pz_tmp, qp, pnew);
                        if (LAST) stb(st.pz_base, qp, pnew);
                    }
                }
                const double gp0 = zs[F_GP * nzp + i] * fdec, gr0 = zs[F_GRHO * nzp + i] * fdec;
                const double xiz = own[i];
                const double div = pa[i] + dz;
                const double pv = -(pb[i] * div) - gp0 * xiz;
                if (valid(i)) {
                    stb(st.p, colb + j * nxb, pv);
                    pcol[j] = pv;
                }
                const double rho = -(r0 * div) - gr0 * xiz;
                double bzv = -rinv * (rho * gz);
                if (src.mode == 3) bzv = bzv + pamp * zpl[j];
                if (act) w0[i] = (i == isrc) ? bzv + srcv : bzv;
            }
        };
        if (segpml) epi(std::true_type{});
        else epi(std::false_type{});
    } else {
        // group 1: flow part of RHS v_z (:1073-1077) to LDS, xi_z update;
        // group 2: RHS v_x flow part (:1067-1071), xi_x update
        const double *D = wave_ptr(f == 1 ? st.dxvz : st.dxvx);
        const double *B0 = wave_ptr(st.a[f == 1 ? 1 : 0]);
        double l1[MS], l2[MS], l3[MS], l4[MS];
#pragma unroll
        for (int i = 0; i < MS; i++) {
            const unsigned r = rowc(a + i);
            l1[i] = ldb(D, colb + r);
            l2[i] = ldb(bg.v0x, gcolb + r);
            l3[i] = ldb(bg.v0z, gcolb + r);
            l4[i] = ldb(B0, colb + r);
        }
        double *const X1 = ca.TN[1], *const X0 = ca.TN[0], *const RV = st.rvx;
#pragma unroll
        for (int i = 0; i < MS; i++) {
            const int j = a + i;
            const unsigned q = colb + j * nxb;
            const bool zero = j == 0 || x == 0 || x == nx - 1;
            const double dzf = Y[i] * zs[F_STR * nzp + i];
            const bool v = valid(i);
            if (f == 1) {
                if (act) wl[c * nzp + a + i] = -(2.0 * l2[i] * l1[i]) - 2.0 * l3[i] * dzf;
                if (v) stb(X1, q, l4[i] + ca.bnext * (zero ? 0.0 : own[i]));
            } else {
                double rvx = -(2.0 * l2[i] * l1[i]) - 2.0 * l3[i] * dzf;
                rvx = zero ? 0.0 : rvx;
                if (v) {
                    stb(RV, q, rvx);
                    stb(X0, q, l4[i] + ca.bnext * (zero ? 0.0 : own[i]));
                }
            }
        }
    }

    // --- d p/dz (ibc 2, :1023) with the z-PML (:1025-1035), RHS v_z, the
    //     zeroed rows/columns (:1083-1093) and the v_z update: group 0
    __syncthreads();
    if (f == 0) {
        double wp[MS + 6];
#pragma unroll
        for (int k = 0; k < MS + 6; k++) wp[k] = pcol[min(max(a - 3 + k, 0), nz - 1)];
        double ep[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (lastseg) {
#pragma unroll
            for (int k = 0; k < 5; k++) ep[k] = pcol[nz - 5 + k];
        }
        seg_rhs<MS>([&](int k) { return wp[k]; }, ep, Y, a, nz, first, lastseg, 2, 0.0, 0.0, cf.c1, cf.c2, cbt);
        auto ltab = [&](int t, int i) { return zt[(5 + t) * 3 * MS + i]; };   // ibc 2 set
        seg_solve<MS>(Y, ltab);
        if (act) { yb3[c * LY + hw2 + 2 * s] = Y[0]; yb3[c * LY + hw2 + 2 * s + 1] = Y[MS - 1]; }
    }
    __syncthreads();
    if (f == 0) {
        auto ltab = [&](int t, int i) { return zt[(5 + t) * 3 * MS + i]; };   // ibc 2 set
        if (act) seg_correct<MS>(Y, s, P, sg.hw, yb3 + c * LY, mv + nband, ltab);
        const int sslot = (LAST && sink.snap) ? out_slot(time, clk.steps, sink.k0, sink.nsnap) : -1;
        const bool snap = sslot >= 0;
        const size_t slot = snap ? ((size_t)b * sink.nsnap + sslot) * sink.nfield : 0;
        auto epi = [&](auto pml_c) {
            constexpr bool PML = decltype(pml_c)::value;
#pragma unroll
            for (int i = 0; i < MS; i++) {
                double lpsi = 0.0, laz = 0.0, lbz = 0.0, lpb = 0.0;
                if (PML) {
                    const int pj = pmlrow(a + i);
                    lpsi = plv(P_T1, pj);
                    lpb = plv(P_B1, pj);
                    laz = plv(P_AZ, pj);
                    lbz = plv(P_BZ, pj);
                }
                const int j = a + i;
                double gpz = Y[i] * zs[F_STR * nzp + i];
                if (PML) {
                    const bool pm = is_pml(j);
                    const double prhs = laz * gpz + lbz * lpsi;
                    gpz = pm ? gpz + lpsi : gpz;
                    const double pnew = lpb + ca.bnext * prhs;
                    if (pm && act) {
                        const unsigned qp = pz1b + pmlrow(j) * nxb;
                        stb(st.pz_tmp, qp, pnew);
                        if (LAST) stb(st.pz_base, qp, pnew);
                    }
                }
                double rvz = -zs[F_RINV * nzp + i] * gpz + (w0[i] + wl[c * nzp + a + i]);
                rvz = (j == 0 || x == 0 || x == nx - 1) ? 0.0 : rvz;
                if (valid(i)) {
                    stb(ca.TN[3], colb + j * nxb, pc[i] + ca.bnext * rvz);
                    if (snap) sink.snap[((slot + (sink.adjoint ? 3 : 5)) * nz + j) * nx + x] = rvz;
                }
            }
        };
        if (segpml) epi(std::true_type{});
        else epi(std::false_type{});
    }
    if (LAST) {
        // FILTER_VARS_2D (physics2d.f90:509-556): group 0 filters v_z, group 1
        // xi_z, group 2 xi_x, each from a window of rows a-5 .. a+MS+4
        __syncthreads();
        const int fi = f == 0 ? 3 : (f == 1 ? 1 : 0);
        const int sslot = sink.snap ? out_slot(time, clk.steps, sink.k0, sink.nsnap) : -1;
        const bool snap = sslot >= 0;
        const size_t slot = snap ? ((size_t)b * sink.nsnap + sslot) * sink.nfield : 0;
        const int oslot = sink.vz_out ? out_slot(time, clk.steps, sink.k0, sink.nout) : -1;
        const bool out = oslot >= 0;
        const int nh = d.nh;
        const bool xin = x >= nh - 1 && x <= nx - nh;
        const double *Fw = wave_ptr((const double *)ca.TN[fi]);
        double *const Aw = wave_ptr(st.a[fi]);
        double w[MS + 10];
#pragma unroll
        for (int q = 0; q < MS + 10; q++) w[q] = ldb(Fw, colb + rowc(a - 5 + q));
        // snapshot field: forward xi_x 0, xi_z 1, v_z 3; adjoint v_z 1
        const int fs = sink.adjoint ? (fi == 3 ? 1 : -1) : fi;
#pragma unroll
        for (int i = 0; i < MS; i++) {
            const int j = a + i;
            const int q = i + 5;
            const double f11 = 0.75390625 * w[q] + 0.5 * 0.41015625 * (w[q - 1] + w[q + 1]) +
                               0.5 * -0.234375 * (w[q - 2] + w[q + 2]) +
                               0.5 * 0.087890625 * (w[q - 3] + w[q + 3]) +
                               0.5 * -0.01953125 * (w[q - 4] + w[q + 4]) +
                               0.5 * 0.001953125 * (w[q - 5] + w[q + 5]);
            const double f2 = 0.5 * (w[q + 1] + w[q - 1]);
            double v = w[q];
            if (xin && j >= 5 && j <= nz - 2) v = (j <= nz - 6) ? f11 : f2;
            if (valid(i)) {
                stb(Aw, colb + j * nxb, v);
                if (fi == 3 && out && j == sink.orow)
                    sink.vz_out[((size_t)b * sink.nout + oslot) * nx + x] = v;
                if (snap && fs >= 0) sink.snap[(slot + fs) * nz * nx + (size_t)j * nx + x] = v;
            }
        }
    }
}

namespace {
int g_lds_max = 64 * 1024;   // dynamic LDS a block may use on this device
}

// Raises every instantiation's dynamic-LDS limit to the device's maximum per
// block (gfx950: 160 KB).
hipError_t colseg_init()
{
    int dev = 0, v = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
        g_lds_max = v;
#define A(MSV)                                                                                          \
    if (hipFuncSetAttribute((const void *)k_colsplit<MSV, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            g_lds_max) != hipSuccess ||                                                 \
        hipFuncSetAttribute((const void *)k_colsplit<MSV, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                            g_lds_max) != hipSuccess) {                                                 \
        (void)hipGetLastError();                                                                        \
        g_lds_max = 64 * 1024;                                                                          \
    }
    SPARC_COLSEG_SIZES(A)
#undef A
    return hipSuccess;
}

}  // namespace sparc

namespace sparc {

bool colsplit_fits(int tc, const Dims &d, const SegTab &sg)
{
    return d.nx % tc == 0 && (tc * sg.P + 63) / 64 <= 2 &&
           colsplit_lds_bytes(tc, sg.P, sg.hw, sg.nzp, d.nzpml, sg.m) <= (size_t)g_lds_max;
}

hipError_t launch_colsplit(int tc, int ms, const Dims &d, const Bg &bg, const DbyCoef &cf, const SegTab &sg,
                           const State &s, const ColArgs &ca, const Src &src, const Clock &c, const Sink &sk,
                           hipStream_t st)
{
    const int G = (tc * sg.P + 63) / 64 * 64;
    if ((tc != 1 && tc != 2 && tc != 4 && tc != 8 && tc != 16) || d.nx % tc || sg.m != ms ||
        sg.P > kSegMax || 3 * G > 384)
        return hipErrorInvalidValue;
    dim3 grid(d.nx / tc, d.S), block(3 * G);
    const size_t shm = colsplit_lds_bytes(tc, sg.P, sg.hw, sg.nzp, d.nzpml, ms);
    if (shm > (size_t)g_lds_max) return hipErrorInvalidValue;
#define L(MSV)                                                                                          \
    if (ms == MSV) {                                                                                    \
        if (ca.last) hipLaunchKernelGGL((k_colsplit<MSV, 1>), grid, block, shm, st, d, bg, cf, sg, s, ca, src, c, sk, tc); \
        else hipLaunchKernelGGL((k_colsplit<MSV, 0>), grid, block, shm, st, d, bg, cf, sg, s, ca, src, c, sk, tc);         \
        return hipGetLastError();                                                                       \
    }
    SPARC_COLSEG_SIZES(L)
#undef L
    return hipErrorInvalidValue;
}

}  // namespace sparc
