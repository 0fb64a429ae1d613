// Memory layout of the MaskedCouplingRQSpline flow, shared by host and device.
//
// The conditioner of every coupling layer is the reference's MLP
// (src/flowMC/resource/model/common.py:68-124, built at
// nf_model/rqSpline.py:427-433): nh hidden layers of widths h[0..nh-1] with
// tanh, then the output layer of d*P units, P = 3K + 1; "linear i" is hidden
// layer i for i < nh and the output layer for i = nh. Its input width is
// in_0 = d and in_i = h[i-1]; its output width is h[i] (i < nh) or d*P.
//
// Master parameters ("theta", FP32, one flat vector) follow the leaves of the
// reference model (rqSpline.py:427-443, common.py:95-107,211-217,243-283),
// layer after layer:
//   [scale, shift, W_0 (h0 x d), b_0 (h0), W_1 (h1 x h0), b_1, ...,
//    W_nh (d*P x h_{nh-1}), b_nh (d*P)]
// followed by data_mean (d), data_cov (d x d), base_mean (d), base_cov (d x d).
// Linear layers use the y = W x + b convention of eqx.nn.Linear.
//
// Layer l couples with mask_l[j] = (j + l) % 2 (rqSpline.py:434): dims with
// mask 1 condition, dims with mask 0 are transformed. Conditioning dim k of
// layer l is 2k + ((l+1)&1); transformed dim k is 2k + (l&1).
//
// Packed MFMA operands (element type T = bf16 or f32) per layer, each a
// "B-fragment" image of an [N x K] matrix (N, K padded) in 32 x KSTEP blocks:
//   wh(0) : W_0 restricted to conditioning columns    [hp(0) x kc]
//   wh(i) : W_i, 0 < i < nh                            [hp(i) x hp(i-1)]
//   w3    : W_nh rows of transformed dims, in chunks of G dims, each chunk
//           padded to a multiple of 32 rows            [N3p x hlp]
//   wht(0): (W_0 restricted)^T                        [ncp x hp(0)]
//   wht(i): W_i^T, 0 < i < nh                          [hp(i-1) x hp(i)]
//   w3t   : chunked W_nh^T                             [hlp x N3p]
//   w3r   : W_nh in the register-spline order          [N3r x hlp]
//   w3rt  : W_nh^T with K in the parameter-gradient order [hlp x N3r]
// (hlp: the last hidden width padded.) With K <= 10 only the register-spline
// images exist (N3p = 0); with K > 10 only the chunked ones (N3r = 0).
// Packed biases (FP32): b_0 .. b_{nh-1} (hp(i) each), b_nh chunked (N3p),
// b_nh in register order (N3r).
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

// most hidden layers of a conditioner (FMC_FLOW_MAX_HIDDEN of include/flowmc.h)
constexpr int FLOW_MAX_HIDDEN = 4;

struct FlowDims {
  int d, L, K;
  int nh;                    // hidden layers of the conditioner, 1..FLOW_MAX_HIDDEN
  int h[FLOW_MAX_HIDDEN];    // their widths (true sizes)
  int kstep;                 // 16 (bf16) or 2 (fp32): K granularity of one MFMA
  int base_dense;            // the base Gaussian's full covariance is used (its Cholesky factor
                             // follows the layers' packed biases: pb_base())
  // derived by finalize() (so device code reads offsets instead of looping
  // over the hidden layers): hsp[i] = sum_{j<i} hp(j); hmm[i] = sum over
  // 0 < j < i of hp(j) hp(j-1); the last and the widest padded hidden width
  // thl[i]: offset of linear i's weight within a layer of theta (after the
  // scale and shift), thl[nh + 1] the layer size
  int hsp[FLOW_MAX_HIDDEN + 1];
  int hmm[FLOW_MAX_HIDDEN + 1];
  int64_t thl[FLOW_MAX_HIDDEN + 2];
  int hl_, hlp_, hmaxp_;
  FMC_HD void finalize() {
    hsp[0] = hmm[0] = 0;
    hmaxp_ = 0;
    for (int i = 0; i < FLOW_MAX_HIDDEN; ++i) {
      const int w = i < nh ? hp(i) : 0;
      hsp[i + 1] = hsp[i] + w;
      hmm[i + 1] = hmm[i] + ((i > 0 && i < nh) ? w * hp(i - 1) : 0);
      hmaxp_ = w > hmaxp_ ? w : hmaxp_;
    }
    hl_ = h[nh - 1];
    hlp_ = hp(nh - 1);
    thl[0] = 2;
    for (int i = 0; i <= FLOW_MAX_HIDDEN; ++i)
      thl[i + 1] = thl[i] + (i <= nh ? (int64_t)lin_out(i) * lin_in(i) + lin_out(i) : 0);
  }
  FMC_HD int P() const { return 3 * K + 1; }
  FMC_HD int hp(int i) const { return rup(h[i], 32); }
  FMC_HD int hl() const { return hl_; }      // last hidden width: the output layer's input
  FMC_HD int hlp() const { return hlp_; }
  FMC_HD int hmaxp() const { return hmaxp_; }
  FMC_HD int hsum_p() const { return hsp[nh]; }  // sum of the padded hidden widths
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
    const int hh = (r >> 2) & 1, s = 16 * b + 4 * (r >> 3) + (r & 3);
    const int k = 2 * q + hh, p = rs_param(s);
    if (k >= dt(l) || p < 0) return -1;
    return trans(l, k) * P() + p;
  }
  // W3 row of parameter-gradient row n = 64q + 16t + 8h + e, or -1
  FMC_HD int w3k_row(int l, int n) const {
    const int q = n >> 6, r = n & 63;
    const int t = r >> 4, hh = (r >> 3) & 1, s = 8 * t + (r & 7);
    const int k = 2 * q + hh, p = rs_param(s);
    if (k >= dt(l) || p < 0) return -1;
    return trans(l, k) * P() + p;
  }
  // W3 row of gradient-store row n (either order)
  FMC_HD int gp_row(int l, int n) const { return reg_spline() ? w3k_row(l, n) : w3_row(l, n); }
  // W3 row of bias slot e = 64q + 32h + s of the register-order bias, or -1
  FMC_HD int b3r_row(int l, int e) const {
    const int q = e >> 6, hh = (e >> 5) & 1, s = e & 31;
    const int k = 2 * q + hh, p = rs_param(s);
    if (k >= dt(l) || p < 0) return -1;
    return trans(l, k) * P() + p;
  }

  // ---- packed operand offsets (elements of T) within the packed buffer
  // padded input width of hidden layer i's image (K of wh(i), N of wht(i))
  FMC_HD int wh_k(int l, int i) const { return i == 0 ? kc(l) : hp(i - 1); }
  FMC_HD int wht_n(int l, int i) const { return i == 0 ? ncp(l) : hp(i - 1); }
  FMC_HD int64_t pk_wh(int l, int i) const {
    return i == 0 ? 0 : (int64_t)hp(0) * kc(l) + hmm[i];
  }
  FMC_HD int64_t pk_w3(int l) const { return pk_wh(l, nh); }
  FMC_HD int64_t pk_wht(int l, int i) const {
    return pk_w3(l) + (int64_t)n3p(l) * hlp() + (i == 0 ? 0 : (int64_t)ncp(l) * hp(0) + hmm[i]);
  }
  FMC_HD int64_t pk_w3t(int l) const { return pk_wht(l, nh); }
  FMC_HD int64_t pk_w3r(int l) const { return pk_w3t(l) + (int64_t)hlp() * n3p(l); }
  FMC_HD int64_t pk_w3rt(int l) const { return pk_w3r(l) + (int64_t)n3r(l) * hlp(); }
  FMC_HD int64_t pk_size(int l) const { return pk_w3rt(l) + (int64_t)hlp() * n3r(l); }
  FMC_HD int64_t pk_layer(int l) const {
    return (int64_t)(l / 2) * (pk_size(0) + pk_size(1)) + ((l & 1) ? pk_size(0) : 0);
  }
  FMC_HD int64_t pk_total() const { return pk_layer(L); }
  // packed biases (FP32)
  FMC_HD int64_t pb_b(int i) const { return hsp[i]; }  // hidden layer i
  FMC_HD int64_t pb_b3(int) const { return pb_b(nh); }  // chunked output-layer bias
  FMC_HD int64_t pb_b3r(int l) const { return pb_b(nh) + n3p(l); }
  FMC_HD int64_t pb_size(int l) const { return pb_b3r(l) + n3r(l); }
  FMC_HD int64_t pb_layer(int l) const {
    return (int64_t)(l / 2) * (pb_size(0) + pb_size(1)) + ((l & 1) ? pb_size(0) : 0);
  }
  FMC_HD int64_t pb_total() const { return pb_layer(L); }
  // chol(base_cov) [d x d] row-major, then sum_j log L_jj (dense base only)
  FMC_HD int64_t pb_base() const { return pb_total(); }
  FMC_HD int64_t pb_alloc() const { return pb_total() + (base_dense ? (int64_t)d * d + 1 : 0); }

  // ---- theta offsets (FP32 master parameters)
  FMC_HD int lin_in(int i) const { return i == 0 ? d : h[i - 1]; }
  FMC_HD int lin_out(int i) const { return i == nh ? d * P() : h[i]; }
  FMC_HD int64_t th_lin(int i) const { return thl[i]; }  // offset of W_i within a layer
  FMC_HD int64_t th_layer_size() const { return th_lin(nh + 1); }
  FMC_HD int64_t th_scale(int l) const { return (int64_t)l * th_layer_size(); }
  FMC_HD int64_t th_shift(int l) const { return th_scale(l) + 1; }
  FMC_HD int64_t th_w(int l, int i) const { return th_scale(l) + th_lin(i); }
  FMC_HD int64_t th_b(int l, int i) const { return th_w(l, i) + (int64_t)lin_out(i) * lin_in(i); }
  FMC_HD int64_t th_w3(int l) const { return th_w(l, nh); }
  FMC_HD int64_t th_b3(int l) const { return th_b(l, nh); }
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
  // Rows: conditioner input c (ncp), hidden activations h_0 .. h_{nh-1},
  // their pre-activation gradients g_0 .. g_{nh-1}, then the spline
  // parameter gradients (n3). Feature sections start at multiples of 32
  // rows. Fragment-blocked: the 32 rows x KSTEP samples of row block rb,
  // batch slice s are one contiguous block of 64 lanes x EPL elements,
  // element (row 32 rb + r, sample KSTEP s + j) at lane r + 32 (j / EPL),
  // position j % EPL, so one MFMA operand fragment along the batch is one
  // coalesced 64-lane read.
  FMC_HD int64_t st_c(int) const { return 0; }
  FMC_HD int64_t st_h(int l, int i) const { return ncp(l) + pb_b(i); }
  FMC_HD int64_t st_gh(int l, int i) const { return ncp(l) + hsum_p() + pb_b(i); }
  FMC_HD int64_t st_gp(int l) const { return ncp(l) + 2 * (int64_t)hsum_p(); }
  FMC_HD int64_t st_rows(int l) const { return st_gp(l) + n3(l); }
};

}  // namespace fmc
