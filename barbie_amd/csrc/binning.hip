// Shape/size binning, Y_lm weights and accumulation (FP64).
// Replaces barbie.c:287-298 (triangle parameters), :301-305 (analytic
// comparison), :343-363 (bins, zero-signal rule, accumulation) and cal_coeff
// (funcs.c:708-745).
//
// Contractions arrive sorted by bin (20-byte records), cut into segments of at most kBinSeg
// records within one bin (so that the largest bins do not set the time: at
// NN=384 one bin holds 26 110 of 2.76 M records).  One workgroup of four
// waves owns one segment: each thread accumulates a strided subset of the
// segment's contractions in registers, each wave reduces its 48 sums in a
// fixed tree order and the four wave sums are added in wave order.  A second
// kernel adds each bin's segment sums in segment order (deterministic).
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace bb {

// 16 orientation weights bs[] and norms norm[] of (px, pz), funcs.c:708-745.
// Contraction to FMA is off here and in rec_coeffs, so that every kernel that
// inlines them computes the same bits.
__device__ __forceinline__ void cal_coeff(double px, double pz, double *bs, double *norm)
{
#pragma clang fp contract(off)
    const double x2 = px * px, x4 = x2 * x2, x6 = x4 * x2;
    const double z2 = pz * pz, z4 = z2 * z2, z6 = z4 * z2;
    const double om = -1. + z2;  // (-1 + pz^2)
    const double u = 1. - z2;    // (1 - pz^2)
    bs[0] = 1.;
    bs[1] = 2. * (3. * z2 - 1.);
    bs[2] = -sqrt(2. / 3.) * px * pz;
    bs[3] = 2. * sqrt(0.6666666666666666) * (-1. + 2. * x2 + z2);
    bs[4] = 8. * (3. - 30. * z2 + 35. * z4);
    bs[5] = (-4. * px * pz * (-3. + 7. * z2)) / sqrt(5.);
    bs[6] = 4. * sqrt(0.4) * (-1 + 2. * x2 + z2) * (-1. + 7. * z2);
    bs[7] = (-4. * px * pz * (-3. + 4. * x2 + 3. * z2)) / sqrt(35.);
    bs[8] = 8. * sqrt(0.05714285714285714) * (8. * x4 + 8. * x2 * om + om * om);
    bs[9] = 16. * (-5. + 105. * z2 - 315. * z4 + 231. * z6);
    bs[10] = -8. * sqrt(0.09523809523809523) * px * pz * (5. - 30. * z2 + 33. * z4);
    bs[11] = (32. * (-1. + 2. * x2 + z2) * (1. - 18. * z2 + 33. * z4)) / sqrt(105.0);
    bs[12] = (-16. * px * pz * (-3. + 4. * x2 + 3. * z2) * (-3. + 11. * z2)) / sqrt(105.0);
    bs[13] = (16. * sqrt(0.2857142857142857) * (-1. + 11. * z2) * (8. * x4 + 8. * x2 * om + om * om)) / 3.;
    bs[14] = (-16 * pz * (16 * x4 * px + 20 * x2 * px * om + 5 * px * om * om)) / (3. * sqrt(77.0));
    bs[15] = (32. * (32. * x6 + 48. * x4 * om + 18. * x2 * om * om + om * om * om)) / sqrt(231.);

    const double a1 = 3. * z2 - 1.;
    const double a4 = 3. - 30. * z2 + 35. * z4;
    const double a5 = -3. * pz + 7. * z2 * pz;
    const double a6 = -1. + 7. * z2;
    const double a9 = -5. + 105. * z2 - 315. * z4 + 231. * z6;
    const double a10 = 5. * pz - 30. * z2 * pz + 33. * z4 * pz;
    const double a11 = 1. - 18. * z2 + 33. * z4;
    const double a12 = -3. * pz + 11. * z2 * pz;
    const double a13 = -1 + 11 * z2;
    const double om2 = om * om, u2 = u * u, u3 = u2 * u;
    norm[0] = 1.;
    norm[1] = a1 * a1;
    norm[2] = z2 * u;
    norm[3] = om2;
    norm[4] = a4 * a4;
    norm[5] = u * a5 * a5;
    norm[6] = om2 * a6 * a6;
    norm[7] = z2 * u3;
    norm[8] = om2 * om2;
    norm[9] = a9 * a9;
    norm[10] = u * a10 * a10;
    norm[11] = om2 * a11 * a11;
    norm[12] = u3 * a12 * a12;
    norm[13] = om2 * om2 * a13 * a13;
    norm[14] = z2 * u3 * u2;
    norm[15] = om2 * om2 * om2;
}

constexpr int kBinThreads = 256;   // four waves per bin

// The field-independent part of one contraction's binning: the triangle
// parameters (barbie.c:287-298; funcs.c:113-126 sort), the weights and norms
// (cal_coeff) and the analytic bispectrum (barbie.c:301-305) before the zero
// rule and the triangle count.
__device__ __forceinline__ void rec_coeffs(const BinRec &R, const double *__restrict__ avg,
                                           const BinParams &bp, double *bs, double *nm,
                                           double &biana)
{
#pragma clang fp contract(off)
    const double a1 = avg[R.r1], a2 = avg[R.r2], a3 = avg[R.r3];
    const double p1 = R.p1, p2 = R.p2, p3 = R.p3;
    double K[3][2] = {{sqrt(a1 * a1 + p1 * p1), p1},
                      {sqrt(a2 * a2 + p2 * p2), p2},
                      {sqrt(a3 * a3 + p3 * p3), p3}};
    // funcs.c:113-126: descending, strict '<'
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2 - i; j++)
            if (K[j][0] < K[j + 1][0]) {
                double s1 = K[j][0], s2 = K[j][1];
                K[j][0] = K[j + 1][0]; K[j][1] = K[j + 1][1];
                K[j + 1][0] = s1; K[j + 1][1] = s2;
            }
    const double T = K[1][0] / K[0][0];
    const double MU = 0.5 * (T + 1. / T - (K[2][0] / K[0][0]) * (K[2][0] / K[1][0]));
    const double pz = K[0][1] / K[0][0];
    const double px = ((K[1][1] / K[1][0]) + MU * pz) / (sqrt(1. - (MU * MU)));
    const double m1 = K[0][1] / K[0][0], m2 = K[1][1] / K[1][0], m3 = K[2][1] / K[2][0];
    const double k0 = bp.Cunits * K[0][0], k1 = bp.Cunits * K[1][0], k2 = bp.Cunits * K[2][0];
    const double P0 = 1.0 / (k0 * k0), P1 = 1.0 / (k1 * k1), P2 = 1.0 / (k2 * k2);
    double B = 2. * bp.ff * (P0 * P1 + P1 * P2 + P2 * P0);
    biana = B * (1 + bp.beta * m1 * m1) * (1 + bp.beta * m2 * m2) * (1 + bp.beta * m3 * m3);
    cal_coeff(px, pz, bs, nm);
}

// One contraction's contributions to its bin's 48 sums (barbie.c:343-363):
// Sv the contraction's sum, Num its triangle count, biana its analytic
// value; the zero rule (barbie.c:355-357) from the fields' nonzero flags.
__device__ __forceinline__ void rec_accumulate(const BinRec &R, double Sv, double Num,
                                               double biana, const double *bs, const double *nm,
                                               const int *__restrict__ nzflag, int Nr,
                                               double *abs_, double *anm, double *abi)
{
#pragma clang fp contract(off)
    const int f1 = R.p1 * Nr + R.r1;   // p1 >= 0
    const int f2 = (R.p2 < 0 ? -R.p2 : R.p2) * Nr + R.r2;
    const int f3 = (R.p3 < 0 ? -R.p3 : R.p3) * Nr + R.r3;
    // The sum is exactly zero iff one of its three fields is identically
    // zero; the fields' zero flags decide it.
    if (!nzflag[f1] || !nzflag[f2] || !nzflag[f3]) {
        biana = 0.;
        Num = 0.;
    }
    const double bn = biana * Num;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        abs_[k] = fma(Sv, bs[k], abs_[k]);
        abi[k] = fma(bn, bs[k], abi[k]);
        anm[k] = fma(Num, nm[k], anm[k]);
    }
}

// Wave sums of the 48 accumulators, then the four waves' sums in wave order,
// to binpart[0..47].
__device__ __forceinline__ void bin_reduce_block(const double *abs_, const double *anm,
                                                 const double *abi, double *__restrict__ binpart)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double part[kBinThreads / 64][48];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const double v0 = wave_sum(abs_[k]), v1 = wave_sum(anm[k]), v2 = wave_sum(abi[k]);
        if (lane == 0) {
            part[wave][k] = v0;
            part[wave][16 + k] = v1;
            part[wave][32 + k] = v2;
        }
    }
    __syncthreads();
    if (tid < 48) {
        double v = part[0][tid];
        for (int w = 1; w < kBinThreads / 64; ++w) v += part[w][tid];
        binpart[tid] = v;
    }
}
// The same for the 16 sums of S bs[ii] alone (the same operations, so the
// same bits as the first 16 of bin_reduce_block).
__device__ __forceinline__ void bin_reduce_block16(const double *abs_,
                                                   double *__restrict__ binpart)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double part[kBinThreads / 64][16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const double v0 = wave_sum(abs_[k]);
        if (lane == 0) part[wave][k] = v0;
    }
    __syncthreads();
    if (tid < 16) {
        double v = part[0][tid];
        for (int w = 1; w < kBinThreads / 64; ++w) v += part[w][tid];
        binpart[tid] = v;
    }
}


// All 48 sums of each segment.  With `skip`, a field whose skip flag is set
// (its norm and analytic sums are taken from the field-independent ones)
// is left out.
__global__ __launch_bounds__(kBinThreads) void k_binning(const int32_t *__restrict__ sg,
                                                const BinRec *__restrict__ recs,
                                                const double *__restrict__ S,
                                                const double *__restrict__ num,
                                                const int *__restrict__ nzflag,
                                                const double *__restrict__ avg, BinParams bp,
                                                double *__restrict__ binpart, BatchStrides st,
                                                const int *__restrict__ skip)
{
    const int sgi = blockIdx.x;   // segment
    const int fb = blockIdx.y;    // field of the batch
    if (skip && skip[fb]) return;   // uniform over the workgroup
    S += fb * st.part;
    nzflag += fb * st.nz;
    binpart += fb * st.binpart;
    const int tid = threadIdx.x;
    double abs_[16], anm[16], abi[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) abs_[k] = anm[k] = abi[k] = 0.0;

    // Software-pipelined over the thread's records: the record and partial
    // sums of the next one are loaded before the current one is binned.
    const int end = sg[sgi + 1];
    int c = sg[sgi] + tid;
    BinRec Rn{};
    double Sn = 0.0;
    if (c < end) {
        Rn = recs[c];
        for (int sp = 0; sp < Rn.nsp; ++sp) Sn += S[(long long)Rn.pos + 256 * sp];
    }
    for (; c < end; c += kBinThreads) {
        const BinRec R = Rn;
        const double Sv = Sn;   // sum of the x-split partials, fixed order
        if (c + kBinThreads < end) {
            Rn = recs[c + kBinThreads];
            Sn = 0.0;
            for (int sp = 0; sp < Rn.nsp; ++sp) Sn += S[(long long)Rn.pos + 256 * sp];
        }
        double bs[16], nm[16], biana;
        rec_coeffs(R, avg, bp, bs, nm, biana);
        rec_accumulate(R, Sv, num[R.rp], biana, bs, nm, nzflag, bp.Nr, abs_, anm, abi);
    }
    bin_reduce_block(abs_, anm, abi, binpart + (long long)sgi * 48);
}

// The 16 sums of S bs[ii] of each segment: k_binning's record loop without
// the norm and analytic sums (which do not depend on the field's values),
// in the same order, so the same bits.  Fewer registers than k_binning:
// more waves per SIMD to hide the scattered partial-sum loads.
//
// Each workgroup also compares its slice of the field's nonzero flags with
// the reference flags and clears match[f] on a difference (match[f] is set
// to 1 by the plane gather, before the flags are raised).
__global__ __launch_bounds__(kBinThreads) void k_binning_s(const int32_t *__restrict__ sg,
                                                  int nseg, const BinRec *__restrict__ recs,
                                                  const double *__restrict__ S,
                                                  const double *__restrict__ avg, BinParams bp,
                                                  const int *__restrict__ nzflag,
                                                  BinStatic bst, BatchStrides st)
{
#pragma clang fp contract(off)
    const int sgi = blockIdx.x;   // segment
    const int fb = blockIdx.y;    // field of the batch
    S += fb * st.part;
    double *binpart_s = bst.binpart_s + fb * st.binpart_s;
    const int tid = threadIdx.x;
    {
        const int per = (bst.nflags + nseg - 1) / nseg, f0 = sgi * per;
        int diff = 0;
        for (int k = tid; k < per && f0 + k < bst.nflags; k += kBinThreads)
            diff |= (nzflag[fb * st.nz + f0 + k] != 0) != (bst.ref[f0 + k] != 0);
        if (__any(diff) && (tid & 63) == 0) atomicAnd(&bst.match[fb], 0);
    }
    double abs_[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) abs_[k] = 0.0;
    const int end = sg[sgi + 1];
    int c = sg[sgi] + tid;
    BinRec Rn{};
    double Sn = 0.0;
    if (c < end) {
        Rn = recs[c];
        for (int sp = 0; sp < Rn.nsp; ++sp) Sn += S[(long long)Rn.pos + 256 * sp];
    }
    for (; c < end; c += kBinThreads) {
        const BinRec R = Rn;
        const double Sv = Sn;   // sum of the x-split partials, fixed order
        if (c + kBinThreads < end) {
            Rn = recs[c + kBinThreads];
            Sn = 0.0;
            for (int sp = 0; sp < Rn.nsp; ++sp) Sn += S[(long long)Rn.pos + 256 * sp];
        }
        double bs[16], nm[16], biana;
        rec_coeffs(R, avg, bp, bs, nm, biana);   // the compiler drops nm and biana
#pragma unroll
        for (int k = 0; k < 16; ++k) abs_[k] = fma(Sv, bs[k], abs_[k]);
    }
    bin_reduce_block16(abs_, binpart_s + (long long)sgi * 16);
}

// The field-independent coefficients of every record, computed once (the
// same code as k_binning): table [33][nrec], rows bs[0..15], nm[0..15], biana.
__global__ __launch_bounds__(256) void k_bin_coeffs(const BinRec *__restrict__ recs, int nrec,
                                                    const double *__restrict__ avg, BinParams bp,
                                                    double *__restrict__ W)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= nrec) return;
    double bs[16], nm[16], biana;
    rec_coeffs(recs[c], avg, bp, bs, nm, biana);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        W[(long long)k * nrec + c] = bs[k];
        W[(long long)(16 + k) * nrec + c] = nm[k];
    }
    W[32LL * nrec + c] = biana;
}

// k_binning with the coefficients read from the table instead of computed:
// the same sums in the same order, so the same bits.  Workgroups of one
// segment for consecutive fields are adjacent (blockIdx.x the field), so the
// segment's coefficients are read from the caches by all but the first.
__global__ __launch_bounds__(kBinThreads) void k_binning_w(const int32_t *__restrict__ sg,
                                                  const BinRec *__restrict__ recs, int nrec,
                                                  const double *__restrict__ W,
                                                  const double *__restrict__ S,
                                                  const double *__restrict__ num,
                                                  const int *__restrict__ nzflag, BinParams bp,
                                                  double *__restrict__ binpart, BatchStrides st,
                                                  const int *__restrict__ skip)
{
    const int fb = blockIdx.x;    // field of the batch
    const int sgi = blockIdx.y;   // segment
    if (skip && skip[fb]) return;   // uniform over the workgroup
    S += fb * st.part;
    nzflag += fb * st.nz;
    binpart += fb * st.binpart;
    const int tid = threadIdx.x;
    double abs_[16], anm[16], abi[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) abs_[k] = anm[k] = abi[k] = 0.0;
    const int end = sg[sgi + 1];
    for (int c = sg[sgi] + tid; c < end; c += kBinThreads) {
        const BinRec R = recs[c];
        double Sv = 0.0;
        for (int sp = 0; sp < R.nsp; ++sp) Sv += S[(long long)R.pos + 256 * sp];
        double bs[16], nm[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            bs[k] = W[(long long)k * nrec + c];
            nm[k] = W[(long long)(16 + k) * nrec + c];
        }
        rec_accumulate(R, Sv, num[R.rp], W[32LL * nrec + c], bs, nm, nzflag, bp.Nr, abs_, anm,
                       abi);
    }
    bin_reduce_block(abs_, anm, abi, binpart + (long long)sgi * 48);
}

// k_binning_s with the weights bs[] read from the table (rows 0..15): the
// same sums in the same order, so the same bits.  Workgroups of one segment
// for consecutive fields are adjacent (blockIdx.x the field).
__global__ __launch_bounds__(kBinThreads) void k_binning_ws(const int32_t *__restrict__ sg,
                                                   int nseg, const BinRec *__restrict__ recs,
                                                   int nrec, const double *__restrict__ W,
                                                   const double *__restrict__ S,
                                                   const int *__restrict__ nzflag, BinStatic bst,
                                                   BatchStrides st)
{
    const int fb = blockIdx.x;    // field of the batch
    const int sgi = blockIdx.y;   // segment
    S += fb * st.part;
    double *binpart_s = bst.binpart_s + fb * st.binpart_s;
    const int tid = threadIdx.x;
    {
        const int per = (bst.nflags + nseg - 1) / nseg, f0 = sgi * per;
        int diff = 0;
        for (int k = tid; k < per && f0 + k < bst.nflags; k += kBinThreads)
            diff |= (nzflag[fb * st.nz + f0 + k] != 0) != (bst.ref[f0 + k] != 0);
        if (__any(diff) && (tid & 63) == 0) atomicAnd(&bst.match[fb], 0);
    }
    double abs_[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) abs_[k] = 0.0;
    const int end = sg[sgi + 1];
    for (int c = sg[sgi] + tid; c < end; c += kBinThreads) {
        const BinRec R = recs[c];
        double Sv = 0.0;
        for (int sp = 0; sp < R.nsp; ++sp) Sv += S[(long long)R.pos + 256 * sp];
#pragma unroll
        for (int k = 0; k < 16; ++k) abs_[k] = fma(Sv, W[(long long)k * nrec + c], abs_[k]);
    }
    bin_reduce_block16(abs_, binpart_s + (long long)sgi * 16);
}

// Bin sums: thread (bin, t < 48) adds the bin's segment sums in order;
// bins without records get zeros.
// With a BinStatic (bs.binpart_s set), the S sums come from binpart_s and
// the norm and analytic sums of a field whose flags match the reference
// from the field-independent bins (bs.sums: [3][nbins * 16], the outputs of
// a full pass with the reference flags).
__global__ __launch_bounds__(256) void k_bin_reduce(const int32_t *__restrict__ bin_sg, int nbins,
                                                    const double *__restrict__ binpart,
                                                    double *__restrict__ out_bs,
                                                    double *__restrict__ out_norm,
                                                    double *__restrict__ out_biana,
                                                    BatchStrides st, BinStatic bst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbins * 48) return;
    const int fb = blockIdx.y;
    binpart += fb * st.binpart;
    const int b = i / 48, t = i - b * 48;
    double v = 0.0;
    if (t < 16 && bst.binpart_s) {
        const double *bp = bst.binpart_s + fb * st.binpart_s;
        for (int sgi = bin_sg[b]; sgi < bin_sg[b + 1]; ++sgi) v += bp[(long long)sgi * 16 + t];
    } else if (t >= 16 && bst.binpart_s && bst.match[fb]) {
        v = bst.sums[(long long)(t >> 4) * nbins * 16 + (long long)b * 16 + (t & 15)];
    } else {
        for (int sgi = bin_sg[b]; sgi < bin_sg[b + 1]; ++sgi) v += binpart[(long long)sgi * 48 + t];
    }
    double *o = t < 16 ? out_bs : t < 32 ? out_norm : out_biana;
    if (o) o[fb * st.bins + (long long)b * 16 + (t & 15)] = v;
}

hipError_t launch_bin_coeffs(const BinRec *recs, int nrec, const double *avg, const BinParams &bp,
                             double *W, hipStream_t s)
{
    if (nrec <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_coeffs, dim3((nrec + 255) / 256), dim3(256), 0, s, recs, nrec, avg,
                       bp, W);
    return hipGetLastError();
}

hipError_t launch_binning(const int32_t *sg, int nseg, const int32_t *bin_sg, const BinRec *recs,
                          const double *S, const double *num, const int *nzflag, const double *avg,
                          const BinParams &bp, double *binpart, double *bins_bs,
                          double *bins_norm, double *bins_biana, const BatchStrides &st, int nf,
                          hipStream_t s, const double *W, int nrec, const BinStatic &bst)
{
    if (bp.nbins <= 0) return hipSuccess;
    if (nf < 1 || nf > 65535) return hipErrorInvalidValue;
    if (nseg > 0 && bst.binpart_s) {
        // S sums for every field; the full pass only for fields whose flags
        // differ from the reference.
        if (nseg > 65535) return hipErrorInvalidValue;
        if (W) {
            hipLaunchKernelGGL(k_binning_ws, dim3(nf, nseg), dim3(kBinThreads), 0, s, sg, nseg, recs,
                               nrec, W, S, nzflag, bst, st);
            hipLaunchKernelGGL(k_binning_w, dim3(nf, nseg), dim3(kBinThreads), 0, s, sg, recs, nrec,
                               W, S, num, nzflag, bp, binpart, st, (const int *)bst.match);
        } else {
            hipLaunchKernelGGL(k_binning_s, dim3(nseg, nf), dim3(kBinThreads), 0, s, sg, nseg, recs,
                               S, avg, bp, nzflag, bst, st);
            hipLaunchKernelGGL(k_binning, dim3(nseg, nf), dim3(kBinThreads), 0, s, sg, recs, S, num,
                               nzflag, avg, bp, binpart, st, (const int *)bst.match);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const int n = bp.nbins * 48;
        hipLaunchKernelGGL(k_bin_reduce, dim3((n + 255) / 256, nf), dim3(256), 0, s, bin_sg,
                           bp.nbins, binpart, bins_bs, bins_norm, bins_biana, st, bst);
        return hipGetLastError();
    }
    if (nseg > 0 && W) {
        if (nseg > 65535) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_binning_w, dim3(nf, nseg), dim3(kBinThreads), 0, s, sg, recs, nrec, W,
                           S, num, nzflag, bp, binpart, st, (const int *)nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    } else if (nseg > 0) {
        hipLaunchKernelGGL(k_binning, dim3(nseg, nf), dim3(kBinThreads), 0, s, sg, recs, S, num,
                           nzflag, avg, bp, binpart, st, (const int *)nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const int n = bp.nbins * 48;
    hipLaunchKernelGGL(k_bin_reduce, dim3((n + 255) / 256, nf), dim3(256), 0, s, bin_sg, bp.nbins,
                       binpart, bins_bs, bins_norm, bins_biana, st, BinStatic{});
    return hipGetLastError();
}

}  // namespace bb
