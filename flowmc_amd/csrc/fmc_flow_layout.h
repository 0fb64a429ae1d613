// Memory layout of the MaskedCouplingRQSpline flow, shared by host and device.
//
// Master parameters ("theta", FP32, one flat vector) follow the leaves of the
// reference model (src/flowMC/resource/model/nf_model/rqSpline.py:427-443,
// common.py:95-107,211-217,243-283), layer after layer:
//   [scale, shift, W1 (h1 x d), b1 (h1), W2 (h2 x h1), b2 (h2),
//    W3 (d*P x h2), b3 (d*P)]                      P = 3K + 1
// followed by data_mean (d), data_cov (d x d), base_mean (d), base_cov (d x d).
// Linear layers use the y = W x + b convention of eqx.nn.Linear.
//
// Layer l couples with mask_l[j] = (j + l) % 2 (rqSpline.py:434): dims with
// mask 1 condition, dims with mask 0 are transformed. Conditioning dim k of
// layer l is 2k + ((l+1)&1); transformed dim k is 2k + (l&1).
//
// Packed MFMA operands (element type T = bf16 or f32) per layer, each a
// "B-fragment" image of an [N x K] matrix (N, K padded) in 32 x KSTEP blocks:
//   w1  : W1 restricted to conditioning columns      [h1p  x kc ]
//   w2  : W2                                         [h2p  x h1p]
//   w3  : W3 rows of transformed dims, in chunks of G dims, each chunk padded
//         to a multiple of 32 rows                    [N3p  x h2p]
//   w1t : (W1 restricted)^T                          [ncp  x h1p]
//   w2t : W2^T                                       [h1p  x h2p]
//   w3t : chunked W3^T                               [h2p  x N3p]
//   w3r : W3 in the register-spline order             [N3r  x h2p]
//   w3rt: W3^T with K in the parameter-gradient order [h2p  x N3r]
// With K <= 10 only the register-spline images exist (N3p = 0); with K > 10
// only the chunked ones (N3r = 0).
// Packed biases (FP32): b1 (h1p), b2 (h2p), b3 chunked (N3p), b3 register
// order (N3r).
//
// Register-spline order: the transformed dims of a layer go in pairs
// q = 0, 1, ...; pair q owns rows 64q .. 64q+63 (two 32-row MFMA blocks b).
// Used as the A operand of a 32x32 MFMA whose B operand is the hidden
// activations of 32 samples, the output lane l holds, for its sample l%32,
// the 16 rows 8(v/4) + 4(l/32) + v%4 of each block (v = 0..15), i.e. the
// 32 "slots" s = 16b + v of dim trans(l, 2q + l/32). Slots 0..9 are the bin
// width logits, 10..19 the height logits, 20..29 the knot-slope logits of
// knots 0..K-1 and slot 30 the slope logit of the last knot K; unused slots
// are zero rows. The register-order bias of pair q is 64 floats: the 32 slots
// of dim 2q, then those of dim 2q+1.
//
// Parameter-gradient order (rows of the gradient store and K of w3rt): row
// 64q + 16t + 8h + e holds slot 8t + e of dim trans(l, 2q + h), so that the
// 8 consecutive K entries an MFMA A-operand lane reads in one K step are 8
// slots of one dim, straight from the registers of the lane that owns it.
#pragma once

#include <stdint.h>

#ifndef FMC_HD
#if defined(__HIPCC__)
#define FMC_HD __host__ __device__
#else
#define FMC_HD
#endif
#endif

namespace fmc {

FMC_HD inline int rup(int v, int m) { return (v + m - 1) / m * m; }

struct FlowDims {
  int d, L, K, h1, h2;   // true sizes
  int kstep;             // 16 (bf16) or 2 (fp32): K granularity of one MFMA
  FMC_HD int P() const { return 3 * K + 1; }
  FMC_HD int h1p() const { return rup(h1, 32); }
  FMC_HD int h2p() const { return rup(h2, 32); }
  FMC_HD int dc(int l) const { return (l & 1) ? (d + 1) / 2 : d / 2; }
  FMC_HD int dt(int l) const { return d - dc(l); }
  FMC_HD int cond(int l, int k) const { return 2 * k + ((l + 1) & 1); }
  FMC_HD int trans(int l, int k) const { return 2 * k + (l & 1); }
  // transformed dims per W3 chunk (chunk output stays <= 128 columns)
  FMC_HD int G() const {
    int g = 128 / P();
    return g < 1 ? 1 : g;
  }
  FMC_HD int ncf() const { return rup(G() * P(), 32); }  // columns of a full chunk
  FMC_HD int n_chunks(int l) const { return (dt(l) + G() - 1) / G(); }
  FMC_HD int chunk_cnt(int l, int c) const {
    int r = dt(l) - c * G();
    return r < G() ? r : G();
  }
  FMC_HD int chunk_nc(int l, int c) const { return rup(chunk_cnt(l, c) * P(), 32); }
  FMC_HD int chunk_off(int l, int c) const { return c * ncf(); }
  FMC_HD int n3p(int l) const {  // chunked rows (K > 10 only)
    int nc = n_chunks(l);
    return (nc == 0 || reg_spline()) ? 0 : (nc - 1) * ncf() + chunk_nc(l, nc - 1);
  }
  FMC_HD int n3(int l) const { return reg_spline() ? n3r(l) : n3p(l); }  // gradient-store rows
  FMC_HD int kc(int l) const { return rup(dc(l) > 0 ? dc(l) : 1, kstep); }
  FMC_HD int ncp(int l) const { return rup(dc(l) > 0 ? dc(l) : 1, 32); }

  // ---- register-spline order of W3 (see the header comment)
  static constexpr int RS_KMAX = 10;
  FMC_HD bool reg_spline() const { return K <= RS_KMAX; }
  FMC_HD int n_pairs(int l) const { return (dt(l) + 1) / 2; }
  FMC_HD int n3r(int l) const { return reg_spline() ? 64 * n_pairs(l) : 0; }
  // parameter index p (0..3K) held in slot s, or -1 for a padding slot
  FMC_HD int rs_param(int s) const {
    if (s < 10) return s < K ? s : -1;
    if (s < 20) return s - 10 < K ? K + s - 10 : -1;
    if (s < 30) return s - 20 < K ? 2 * K + s - 20 : -1;
    return s == 30 ? 3 * K : -1;
  }
  // W3 row (within the full d*P rows) of register-order row n, or -1
  FMC_HD int w3r_row(int l, int n) const {
    const int q = n >> 6, b = (n >> 5) & 1, r = n & 31;
    const int h = (r >> 2) & 1, s = 16 * b + 4 * (r >> 3) + (r & 3);
    const int k = 2 * q + h, p = rs_param(s);
    if (k >= dt(l) || p < 0) return -1;
    return trans(l, k) * P() + p;
  }
  // W3 row of parameter-gradient row n = 64q + 16t + 8h + e, or -1
  FMC_HD int w3k_row(int l, int n) const {
    const int q = n >> 6, r = n & 63;
    const int t = r >> 4, h = (r >> 3) & 1, s = 8 * t + (r & 7);
    const int k = 2 * q + h, p = rs_param(s);
    if (k >= dt(l) || p < 0) return -1;
    return trans(l, k) * P() + p;
  }
  // W3 row of gradient-store row n (either order)
  FMC_HD int gp_row(int l, int n) const { return reg_spline() ? w3k_row(l, n) : w3_row(l, n); }
  // W3 row of bias slot e = 64q + 32h + s of the register-order bias, or -1
  FMC_HD int b3r_row(int l, int e) const {
    const int q = e >> 6, h = (e >> 5) & 1, s = e & 31;
    const int k = 2 * q + h, p = rs_param(s);
    if (k >= dt(l) || p < 0) return -1;
    return trans(l, k) * P() + p;
  }

  // ---- packed operand offsets (elements of T) within the packed buffer
  FMC_HD int64_t pk_w1(int) const { return 0; }
  FMC_HD int64_t pk_w2(int l) const { return (int64_t)h1p() * kc(l); }
  FMC_HD int64_t pk_w3(int l) const { return pk_w2(l) + (int64_t)h2p() * h1p(); }
  FMC_HD int64_t pk_w1t(int l) const { return pk_w3(l) + (int64_t)n3p(l) * h2p(); }
  FMC_HD int64_t pk_w2t(int l) const { return pk_w1t(l) + (int64_t)ncp(l) * h1p(); }
  FMC_HD int64_t pk_w3t(int l) const { return pk_w2t(l) + (int64_t)h1p() * h2p(); }
  FMC_HD int64_t pk_w3r(int l) const { return pk_w3t(l) + (int64_t)h2p() * n3p(l); }
  FMC_HD int64_t pk_w3rt(int l) const { return pk_w3r(l) + (int64_t)n3r(l) * h2p(); }
  FMC_HD int64_t pk_size(int l) const { return pk_w3rt(l) + (int64_t)h2p() * n3r(l); }
  FMC_HD int64_t pk_layer(int l) const {
    return (int64_t)(l / 2) * (pk_size(0) + pk_size(1)) + ((l & 1) ? pk_size(0) : 0);
  }
  FMC_HD int64_t pk_total() const { return pk_layer(L); }
  // packed biases (FP32)
  FMC_HD int64_t pb_b3r(int l) const { return (int64_t)h1p() + h2p() + n3p(l); }
  FMC_HD int64_t pb_size(int l) const { return pb_b3r(l) + n3r(l); }
  FMC_HD int64_t pb_layer(int l) const {
    return (int64_t)(l / 2) * (pb_size(0) + pb_size(1)) + ((l & 1) ? pb_size(0) : 0);
  }
  FMC_HD int64_t pb_total() const { return pb_layer(L); }

  // ---- theta offsets (FP32 master parameters)
  FMC_HD int64_t th_layer_size() const {
    return 2 + (int64_t)h1 * d + h1 + (int64_t)h2 * h1 + h2 + (int64_t)d * P() * h2 + (int64_t)d * P();
  }
  FMC_HD int64_t th_scale(int l) const { return (int64_t)l * th_layer_size(); }
  FMC_HD int64_t th_shift(int l) const { return th_scale(l) + 1; }
  FMC_HD int64_t th_w1(int l) const { return th_scale(l) + 2; }
  FMC_HD int64_t th_b1(int l) const { return th_w1(l) + (int64_t)h1 * d; }
  FMC_HD int64_t th_w2(int l) const { return th_b1(l) + h1; }
  FMC_HD int64_t th_b2(int l) const { return th_w2(l) + (int64_t)h2 * h1; }
  FMC_HD int64_t th_w3(int l) const { return th_b2(l) + h2; }
  FMC_HD int64_t th_b3(int l) const { return th_w3(l) + (int64_t)d * P() * h2; }
  FMC_HD int64_t th_data_mean() const { return (int64_t)L * th_layer_size(); }
  FMC_HD int64_t th_data_cov() const { return th_data_mean() + d; }
  FMC_HD int64_t th_base_mean() const { return th_data_cov() + (int64_t)d * d; }
  FMC_HD int64_t th_base_cov() const { return th_base_mean() + d; }
  FMC_HD int64_t th_total() const { return th_base_cov() + (int64_t)d * d; }

  // W3 row (within the full d*P rows) of chunk-padded row n of layer l, or -1
  FMC_HD int w3_row(int l, int n) const {
    int c = n / ncf();
    int r = n - c * ncf();
    if (c >= n_chunks(l) || r >= chunk_cnt(l, c) * P()) return -1;
    int jj = r / P(), p = r - jj * P();
    return trans(l, c * G() + jj) * P() + p;
  }

  // ---- activation / gradient store of one backward layer (elements of T).
  // Feature sections start at multiples of 32 rows. Fragment-blocked: the
  // 32 rows x KSTEP samples of row block rb, batch slice s are one contiguous
  // block of 64 lanes x EPL elements, element (row 32 rb + r, sample
  // KSTEP s + j) at lane r + 32 (j / EPL), position j % EPL, so one MFMA
  // operand fragment along the batch is one coalesced 64-lane read.
  FMC_HD int64_t st_c(int) const { return 0; }
  FMC_HD int64_t st_h1(int l) const { return ncp(l); }
  FMC_HD int64_t st_h2(int l) const { return st_h1(l) + h1p(); }
  FMC_HD int64_t st_gp1(int l) const { return st_h2(l) + h2p(); }
  FMC_HD int64_t st_gp2(int l) const { return st_gp1(l) + h1p(); }
  FMC_HD int64_t st_gp(int l) const { return st_gp2(l) + h2p(); }
  FMC_HD int64_t st_rows(int l) const { return st_gp(l) + n3(l); }
};

}  // namespace fmc
