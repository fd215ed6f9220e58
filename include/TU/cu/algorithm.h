/*!
  \file   algorithm.h
  \brief  The parts of the reference's TU/cu/algorithm.h on the
          FIRFilter2 / FIRGaussianConvolver2 path and the pyramid step after
          it: BlockTraits, divUp and subsample().

  Host-only C++ over the C ABI of libtucuda_amd.so (include/tucu.h).
*/
#pragma once

#include "TU/cu/Array++.h"
#include "TU/cu/functional.h"

#include <iterator>

namespace TU
{
namespace cu
{
//! Thread-block shape of the reference's kernels (algorithm.h:58-67); kept
//! as a template parameter for source compatibility.  The gfx950 kernels
//! choose their own tiling.
template <size_t BLOCK_DIM_X = 32, size_t BLOCK_DIM_Y = 16>
struct BlockTraits
{
    constexpr static size_t BlockDimX = BLOCK_DIM_X;
    constexpr static size_t BlockDimY = BLOCK_DIM_Y;
    constexpr static size_t BlockDim  = BlockDimX * BlockDimY;
};

inline size_t
divUp(size_t dim, size_t blockDim)
{
    return (dim + blockDim - 1) / blockDim;
}

//! algorithm.h:183-242: halves rows [in, ie) in both directions into out,
//! out[y][x] = op at input pixel (2y, 2x) for y < nrow/2, x < ncol/2, with
//! the operator's border rule (0 where its stencil leaves the image).
//! Fewer than 2 rows or columns: no-op.  Operators other than identity1x1
//! need a floating-point input.  Asynchronous with respect to the host; the
//! trailing stream (a hipStream_t, default the null stream) is an extension.
//! Exactly nrow/2 x ncol/2 pixels are written (the reference's grid can
//! write one row / column more for odd extents).
template <class BLOCK_TRAITS = BlockTraits<>, class IN, class OUT, class OP> void
subsample(IN in, IN ie, OUT out, OP, void* stream = nullptr)
{
    const tucu_image src = make_image(in, ie);
    const tucu_image dst = make_out_image(out);
    check(tucu_subsample(&src, &dst, OP::tucu_op, stream));
}

}   // namespace cu
}   // namespace TU
