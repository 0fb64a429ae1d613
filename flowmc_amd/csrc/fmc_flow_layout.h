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
// Packed biases (FP32): b1 (h1p), b2 (h2p), b3 chunked (N3p).
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
  FMC_HD int n3p(int l) const {
    int nc = n_chunks(l);
    return nc == 0 ? 0 : (nc - 1) * ncf() + chunk_nc(l, nc - 1);
  }
  FMC_HD int kc(int l) const { return rup(dc(l) > 0 ? dc(l) : 1, kstep); }
  FMC_HD int ncp(int l) const { return rup(dc(l) > 0 ? dc(l) : 1, 32); }

  // ---- packed operand offsets (elements of T) within the packed buffer
  FMC_HD int64_t pk_w1(int) const { return 0; }
  FMC_HD int64_t pk_w2(int l) const { return (int64_t)h1p() * kc(l); }
  FMC_HD int64_t pk_w3(int l) const { return pk_w2(l) + (int64_t)h2p() * h1p(); }
  FMC_HD int64_t pk_w1t(int l) const { return pk_w3(l) + (int64_t)n3p(l) * h2p(); }
  FMC_HD int64_t pk_w2t(int l) const { return pk_w1t(l) + (int64_t)ncp(l) * h1p(); }
  FMC_HD int64_t pk_w3t(int l) const { return pk_w2t(l) + (int64_t)h1p() * h2p(); }
  FMC_HD int64_t pk_size(int l) const { return pk_w3t(l) + (int64_t)h2p() * n3p(l); }
  FMC_HD int64_t pk_layer(int l) const {
    return (int64_t)(l / 2) * (pk_size(0) + pk_size(1)) + ((l & 1) ? pk_size(0) : 0);
  }
  FMC_HD int64_t pk_total() const { return pk_layer(L); }
  // packed biases (FP32)
  FMC_HD int64_t pb_size(int l) const { return (int64_t)h1p() + h2p() + n3p(l); }
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

  // ---- feature-major activation store of one backward layer ([features, ld] of T)
  FMC_HD int64_t st_c(int) const { return 0; }
  FMC_HD int64_t st_h1(int l) const { return kc(l); }
  FMC_HD int64_t st_h2(int l) const { return st_h1(l) + h1p(); }
  FMC_HD int64_t st_gp1(int l) const { return st_h2(l) + h2p(); }
  FMC_HD int64_t st_gp2(int l) const { return st_gp1(l) + h1p(); }
  FMC_HD int64_t st_gp(int l) const { return st_gp2(l) + h2p(); }
  FMC_HD int64_t st_rows(int l) const { return st_gp(l) + n3p(l); }
};

}  // namespace fmc
