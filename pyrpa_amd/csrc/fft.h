// Batched multi-dimensional complex FFT over the (k-mesh x imaginary-time) grid.
//
// Grid layout: element (x,y,z,t) at x + Nx*(y + Ny*(z + Nz*t)), x fastest, the
// same as the reference's complex(0:Nx-1,0:Ny-1,0:Nz-1,2*Nw) work arrays
// (fchi.f90:529).  Conventions of the reference's FFT wrapper
// (fmod.f90:67-160): sign -1 unnormalised forward, sign +1 backward divided by
// the number of transformed points.
#pragma once
#include "gmod_common.h"

namespace gmod {

struct FFTPassDesc {
  int n = 1;                // transform length
  long n_inner = 1, n_mid = 1, n_outer = 1;
  long in_mid = 0, in_outer = 0, in_elem = 1;
  long out_mid = 0, out_outer = 0, out_elem = 1;
  int twn = 0;              // four-step twiddle: out[k] *= w_twn^(k*mid)
  int src = 0, dst = 0;     // 0 = caller buffer, 1 = scratch
  int lines = 1;            // lines per work-group
  long tw_off = 0;          // offset of the length-n table in the twiddle pool
  long tw4_off = 0;         // offset of the length-twn table
};

template <typename R>
class FFTPlan {
 public:
  // axes_mask: bit a set -> transform axis a (0:x 1:y 2:z 3:t)
  FFTPlan(int nx, int ny, int nz, int m, int axes_mask = 15);
  ~FFTPlan();
  // In place on `buf` (scratch `tmp` of the same size).  sign=-1 forward,
  // sign=+1 backward; `scale` multiplies the result (1/N for the normalised
  // backward transform).
  void exec(cpx<R>* buf, cpx<R>* tmp, int sign, R scale, hipStream_t s) const;
  long size() const { return (long)dims_[0] * dims_[1] * dims_[2] * dims_[3]; }
  long transformed_points() const;
  int num_passes() const { return (int)passes_.size(); }

 private:
  int dims_[4];
  int mask_;
  std::vector<FFTPassDesc> passes_;
  cpx<R>* tw_[2] = {nullptr, nullptr};  // [0]: sign -1 tables, [1]: sign +1 tables
  long tw_len_ = 0;
};

// Cached plans keyed by shape and precision.
template <typename R>
const FFTPlan<R>& fft_plan(int nx, int ny, int nz, int m, int axes_mask = 15);

}  // namespace gmod
