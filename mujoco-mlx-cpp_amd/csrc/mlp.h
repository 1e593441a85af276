// Learned warm start: the MLP object shared by the C API (mlp.hip) and the
// batched engine (engine.hip).
//
// Network: nin -> nh -> nh -> nout, SiLU on the hidden layers, y = x W + b.
// Parameters are FP32 master copies with BF16 shadows used by the MFMA GEMMs.
// Internally the input and output widths are padded to multiples of 32 (the
// GEMM K step); padded weights are zero and receive zero gradient.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

struct MjmlxMLP {
  int nin = 0, nh = 0, nout = 0;
  int ip = 0, op = 0;  // padded input / output widths
  hipStream_t stream = nullptr;
  // parameter blocks W1 [ip x nh], b1, W2 [nh x nh], b2, W3 [nh x op], b3 [op]
  size_t off[6] = {}, len[6] = {}, nparam = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* pb = nullptr;
  long step = 0;
  // normalisation: x_n = (x - in_mean) * in_istd;  y = y_n * out_std + out_mean
  float *in_mean = nullptr, *in_istd = nullptr, *out_mean = nullptr, *out_std = nullptr;
  std::vector<float> h_in_mean, h_in_istd, h_out_mean, h_out_std;
  // activations for up to `cap` rows
  int cap = 0;
  uint16_t *xb = nullptr, *h1b = nullptr, *h2b = nullptr, *d3b = nullptr, *d2b = nullptr, *d1b = nullptr;
  float *z1 = nullptr, *z2 = nullptr, *out = nullptr, *tgt = nullptr, *dout = nullptr, *dh2 = nullptr,
        *dh1 = nullptr, *d2f = nullptr, *d1f = nullptr, *rowloss = nullptr;
  int* d_idx = nullptr;
  long idx_cap = 0;
  // split-K partial tiles of the weight gradients and partial column sums
  float* part = nullptr;
  size_t part_cap = 0;
  // device-resident training set: raw (unnormalised) rows
  float *ds_x = nullptr, *ds_y = nullptr;
  long ds_cap = 0, ds_n = 0;

  uint16_t* Wb(int l) const { return pb + off[2 * l]; }
  float* bias(int l) const { return p + off[2 * l + 1]; }
  float* gW(int l) const { return g + off[2 * l]; }
  float* gb(int l) const { return g + off[2 * l + 1]; }
  void ensure(int rows);
};

namespace mjmlp {

// xb[r] = bf16((x - mean) * istd) for the rows of a raw feature matrix
// [rows][nf] whose columns [ctrl_off, ctrl_off + nu) are replaced by ctrl[r]
// (ctrl may be null: zeros)
void pack_features(MjmlxMLP* n, const float* raw, const float* ctrl, int ctrl_off, int nu, int rows, hipStream_t s);
// pred[r][0..nout) = de-normalised network output for the packed rows in xb,
// plus base[r * base_ld + c] when base is set (residual networks)
void predict(MjmlxMLP* n, int rows, float* pred, hipStream_t s, const float* base = nullptr, int base_ld = 0);
// appends rows (raw features with ctrl inserted, targets minus base[r * base_ld
// + c] when base is set) to the training set; returns the number appended
// (limited by the remaining capacity)
long record(MjmlxMLP* n, const float* raw, const float* ctrl, int ctrl_off, int nu, const float* target, int rows,
            hipStream_t s, const float* base = nullptr, int base_ld = 0);

}  // namespace mjmlp
