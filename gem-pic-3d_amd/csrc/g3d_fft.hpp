// Batched in-place complex FFT of the transverse planes on gfx950.
//
// A plane holds ny rows of nx contiguous points (index j*nx + ix).  Two kernels: rows (x transforms of a tile
// of whole rows per workgroup) and columns (y transforms of a tile of nc adjacent columns per workgroup, loaded
// row by row so that every global access is a contiguous run of nc complex values).  Inside a workgroup the data
// sits in LDS and a radix-2 decimation-in-time transform runs in place after a bit-reversed load, with the
// twiddles exp(-2 pi i k/n) of the reference's own FFT (src/fft.hpp:156-163) staged in LDS; the butterfly is
// u + v w, u - v w as in src/fft.hpp:74-85.  Forward has the sign -1; the inverse is scaled by 1/(nx ny) (2D)
// or 1/nx (rows only) in the last pass (an exact power of two, so the place of the scaling does not round).
#pragma once
#include "g3d_common.hpp"

namespace g3d {

struct FftPlan {
    int nx = 0, ny = 0, two_d = 1, lx = 0, ly = 0;
    cx* twx = nullptr;      // nx/2 twiddles (device)
    cx* twy = nullptr;      // ny/2 twiddles (device)
};

// plan for nx, ny (powers of two 2..1024); throws std::runtime_error on bad sizes
void fft_plan_init(FftPlan& p, int nx, int ny, bool two_d);
void fft_plan_free(FftPlan& p);
// count planes starting at data, in place, on the library stream
void fft_run(const FftPlan& p, cx* data, int count, bool inverse);
// the same in FP32 (FP32-mixed form)
void fft_run_f(const FftPlan& p, cxf* data, int count, bool inverse);
// rows only (x transforms) / columns only (y transforms) of count planes; scale multiplies the result
void fft_rows(const FftPlan& p, cx* data, long nrows, bool inverse, double scale);
void fft_cols(const FftPlan& p, cx* data, int count, bool inverse, double scale);

}  // namespace g3d

// the plan behind the C ABI handle g3d_fft (include/g3d.h)
struct g3d_fft { g3d::FftPlan p; };
