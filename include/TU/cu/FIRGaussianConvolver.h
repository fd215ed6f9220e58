/*!
  \file   FIRGaussianConvolver.h
  \brief  TU::cu::FIRGaussianConvolver2: Gaussian smoothing and its first and
          second derivatives -- the drop-in for the reference's
          TU/cu/FIRGaussianConvolver.h:56-242.

  The six methods keep their (horizontal, vertical) lobe pairs and
  signatures.  convolve6() is an extension that produces any subset of the
  six outputs from one fused launch (one read of the input).
*/
#pragma once

#include "TU/cu/FIRFilter.h"

#include <array>

namespace TU
{
namespace cu
{
/************************************************************************
*  class FIRGaussianConvolver2<T>                                       *
************************************************************************/
template <class T = float>
class FIRGaussianConvolver2 : public FIRFilter2<T>
{
  private:
    using super = FIRFilter2<T>;

  public:
    enum Method { SMOOTH = 0, DIFF_H, DIFF_V, DIFF_HH, DIFF_HV, DIFF_VV };

    //! Implicit, as the reference's (FIRGaussianConvolver.h:63):
    //! FIRGaussianConvolver2<float> g = 2.0f; compiles.
    FIRGaussianConvolver2(float sigma = 1.0, int rule = TUCU_RULE_REFERENCE)
        : _rule(rule)
    {
        initialize(sigma);
    }

    float sigma() const { return _sigma; }

    //! FIRGaussianConvolver.h:114-121.
    FIRGaussianConvolver2& initialize(float sigma)
    {
        _sigma = sigma;
        initialize(_sigma, _lobe0, _lobe1, _lobe2, _rule);
        return *this;
    }

    //! FIRGaussianConvolver.h:123-170: lobes of the Gaussian and of its
    //! first and second derivatives, lengths by the 1 %-of-peak rule.
    template <class LOBE>
    static void initialize(float sigma, LOBE& lobe0, LOBE& lobe1, LOBE& lobe2,
                           int rule = TUCU_RULE_REFERENCE)
    {
        float l[3][17];
        int   n[3];
        check(tucu_gauss2_lobes(sigma, rule, l[0], &n[0], l[1], &n[1], l[2], &n[2]));
        LOBE* dst[3] = {&lobe0, &lobe1, &lobe2};
        for (int k = 0; k < 3; ++k) {
            dst[k]->resize(size_t(n[k]));
            for (int i = 0; i < n[k]; ++i) (*dst[k])[i] = l[k][i];
        }
    }

    template <class IN, class OUT>
    void smooth(IN in, IN ie, OUT out, bool shift = true)   { apply(SMOOTH, in, ie, out, shift); }   // :178-182
    template <class IN, class OUT>
    void diffH(IN in, IN ie, OUT out, bool shift = true)    { apply(DIFF_H, in, ie, out, shift); }   // :190-194
    template <class IN, class OUT>
    void diffV(IN in, IN ie, OUT out, bool shift = true)    { apply(DIFF_V, in, ie, out, shift); }   // :202-206
    template <class IN, class OUT>
    void diffHH(IN in, IN ie, OUT out, bool shift = true)   { apply(DIFF_HH, in, ie, out, shift); }  // :214-218
    template <class IN, class OUT>
    void diffHV(IN in, IN ie, OUT out, bool shift = true)   { apply(DIFF_HV, in, ie, out, shift); }  // :226-230
    template <class IN, class OUT>
    void diffVV(IN in, IN ie, OUT out, bool shift = true)   { apply(DIFF_VV, in, ie, out, shift); }  // :238-242

    //! Extension: the outputs selected by mask (bit k = Method k) from one
    //! fused launch; outs[k] is where method k writes.  Every output equals
    //! what the corresponding single method writes.
    template <class IN, class OUT>
    void convolve6(IN in, IN ie, const std::array<OUT, 6>& outs, unsigned mask = TUCU_ALL6,
                   bool shift = true) const
    {
        const tucu_image src = make_image(in, ie);
        tucu_image dst[6] = {};
        for (int k = 0; k < 6; ++k)
            if (mask >> k & 1u) dst[k] = make_out_image(outs[k]);
        check(tucu_gauss2_convolve(_lobe0.data(), int(_lobe0.size()), _lobe1.data(), int(_lobe1.size()),
                                   _lobe2.data(), int(_lobe2.size()), &src, dst, mask,
                                   shift ? 1 : 0, super::stream()));
    }

    //! Extension: the pyramid step in one launch.  out (nrow/2 x ncol/2)
    //! receives what smooth() into a zero-filled image followed by
    //! subsample(..., identity1x1()) gives, without the full-size
    //! intermediate.
    template <class IN, class OUT>
    void smoothSubsample(IN in, IN ie, OUT out, bool shift = true) const
    {
        const tucu_image src = make_image(in, ie);
        const tucu_image dst = make_out_image(out);
        check(tucu_gauss2_pyramid(_lobe0.data(), int(_lobe0.size()), &src, &dst, shift ? 1 : 0,
                                  super::stream()));
    }

  private:
    template <class IN, class OUT>
    void apply(Method m, IN in, IN ie, OUT out, bool shift)
    {
        if (_rule != TUCU_RULE_REFERENCE) {         // lobes outside FIRFilter2's size classes
            std::array<OUT, 6> outs;
            outs.fill(out);
            convolve6(in, ie, outs, 1u << m, shift);
            return;
        }
        static const int HV[6][2] = {{0, 0}, {1, 0}, {0, 1}, {2, 0}, {1, 1}, {0, 2}};
        const std::vector<float>* lobes[3] = {&_lobe0, &_lobe1, &_lobe2};
        super::initialize(*lobes[HV[m][0]], *lobes[HV[m][1]]).convolve(in, ie, out, shift);
    }

    float              _sigma;
    int                _rule;
    std::vector<float> _lobe0;      //!< smoothing lobe
    std::vector<float> _lobe1;      //!< first-derivative lobe
    std::vector<float> _lobe2;      //!< second-derivative lobe
};

}   // namespace cu
}   // namespace TU
