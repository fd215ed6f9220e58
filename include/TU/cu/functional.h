/*!
  \file   functional.h
  \brief  The stencil operators of the reference's TU/cu/functional.h:58-256
          that cu::subsample() takes, as tag types for the drop-in.

  Each type carries the reference's OperatorSizeY / OperatorSizeX and the
  TUCU_OP_* code the C ABI evaluates on the GPU (tucu_subsample); the
  arithmetic itself, with each operator's border rule (0 where the stencil
  would leave the image), lives in libtucuda_amd.so.  The logical_* and
  extremal3x3 operators, which serve other algorithms of the reference,
  are not provided.
*/
#pragma once

#include "tucu.h"

#include <cstddef>

namespace TU
{
namespace cu
{
template <size_t OPERATOR_SIZE_Y, size_t OPERATOR_SIZE_X, int CODE>
struct OperatorTraits
{
    constexpr static size_t OperatorSizeY = OPERATOR_SIZE_Y;
    constexpr static size_t OperatorSizeX = OPERATOR_SIZE_X;
    constexpr static int    tucu_op       = CODE;     //!< TUCU_OP_* of the C ABI
};

struct identity1x1  : OperatorTraits<1, 1, TUCU_OP_IDENTITY1X1>  {};   // functional.h:58-65
struct diffH1x3     : OperatorTraits<1, 3, TUCU_OP_DIFFH1X3>     {};   // :68-81
struct diffV3x1     : OperatorTraits<3, 1, TUCU_OP_DIFFV3X1>     {};   // :84-98
struct diffHH1x3    : OperatorTraits<1, 3, TUCU_OP_DIFFHH1X3>    {};   // :101-114
struct diffVV3x1    : OperatorTraits<3, 1, TUCU_OP_DIFFVV3X1>    {};   // :117-130
struct diffHV3x3    : OperatorTraits<3, 3, TUCU_OP_DIFFHV3X3>    {};   // :133-148
struct sobelH3x3    : OperatorTraits<3, 3, TUCU_OP_SOBELH3X3>    {};   // :151-167
struct sobelV3x3    : OperatorTraits<3, 3, TUCU_OP_SOBELV3X3>    {};   // :170-186
struct sobelAbs3x3  : OperatorTraits<3, 3, TUCU_OP_SOBELABS3X3>  {};   // :189-198
struct laplacian3x3 : OperatorTraits<3, 3, TUCU_OP_LAPLACIAN3X3> {};   // :201-210
struct det3x3       : OperatorTraits<3, 3, TUCU_OP_DET3X3>       {};   // :213-226
struct gaussian5x5  : OperatorTraits<5, 5, TUCU_OP_GAUSSIAN5X5>  {};   // :229-256

}   // namespace cu
}   // namespace TU
