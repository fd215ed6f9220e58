/*!
  \file   FIRFilter.h
  \brief  TU::cu::FIRFilter2: separable 2-D FIR filter with symmetric (odd
          length) or antisymmetric (even length) half-kernels -- the
          drop-in for the reference's TU/cu/FIRFilter.h:18-52.

  Same class, method names, argument meanings and exceptions; the work runs
  in libtucuda_amd.so's fused gfx950 kernel through the C ABI.  Differences
  a caller can observe: coefficients travel with each launch (no process-
  global constant bank, so filters on different threads do not interfere),
  no intermediate buffer is allocated, and setStream() selects the HIP
  stream (default: the null stream, as in the reference).
*/
#pragma once

#include "TU/cu/algorithm.h"

#include <iterator>
#include <vector>

namespace TU
{
namespace cu
{
/************************************************************************
*  class FIRFilter2<T, BLOCK_TRAITS>                                    *
************************************************************************/
template <class T = float, class BLOCK_TRAITS = BlockTraits<> >
class FIRFilter2 : public BLOCK_TRAITS
{
    static_assert(std::is_same<T, float>::value, "the intermediate type is float");

  public:
    using value_type = T;

    using BLOCK_TRAITS::BlockDimX;
    using BLOCK_TRAITS::BlockDimY;
    constexpr static size_t LobeSizeMax = 17;

  public:
    FIRFilter2() : _stream(nullptr) {}

    //! FIRFilter.h:245-261.  Throws std::runtime_error when a lobe is longer
    //! than 17; lengths are checked against {2,3,4,5,8,9,16,17} at convolve().
    template <class LOBE>
    FIRFilter2& initialize(const LOBE& lobeH, const LOBE& lobeV)
    {
        check(tucu_fir2_check(int(std::size(lobeH)), int(std::size(lobeV))));
        _lobeH.assign(std::data(lobeH), std::data(lobeH) + std::size(lobeH));
        _lobeV.assign(std::data(lobeV), std::data(lobeV) + std::size(lobeV));
        return *this;
    }

    //! FIRFilter.h:298-377: filters rows [ib, ie) into out.  Only the
    //! (nrow - 2 offsetV()) x (ncol - 2 offsetH()) valid region is written,
    //! at (offsetV(), offsetH()) when shift is true; too-small images are a
    //! no-op.  Asynchronous with respect to the host.
    template <class IN, class OUT>
    void convolve(IN ib, IN ie, OUT out, bool shift = true) const
    {
        const tucu_image in = make_image(ib, ie);
        const tucu_image o  = make_out_image(out);
        check(tucu_fir2_convolve(_lobeH.data(), int(_lobeH.size()), _lobeV.data(), int(_lobeV.size()),
                                 &in, &o, shift ? 1 : 0, _stream));
    }

    size_t outSizeH(size_t inSize) const { return inSize - 2 * offsetH(); }   // FIRFilter.h:37
    size_t outSizeV(size_t inSize) const { return inSize - 2 * offsetV(); }   // :38
    size_t offsetH()               const { return _lobeH.size() & ~0x1; }     // :39
    size_t offsetV()               const { return _lobeV.size() & ~0x1; }     // :40

    //! Extension: the HIP stream (a hipStream_t) later calls enqueue on.
    void   setStream(void* stream)       { _stream = stream; }
    void*  stream()                const { return _stream; }

  protected:
    const std::vector<float>& lobeH() const { return _lobeH; }
    const std::vector<float>& lobeV() const { return _lobeV; }

  private:
    std::vector<float> _lobeH;
    std::vector<float> _lobeV;
    void*              _stream;
};

}   // namespace cu
}   // namespace TU
