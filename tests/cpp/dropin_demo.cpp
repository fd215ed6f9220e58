// dropin_demo.cpp -- the reference's gaussTest / filterTest harness shape
// (Testprog/gaussTest/main.cu:56-148, Testprog/filterTest/main.cu:54-145)
// written against the drop-in headers, so that the drop-in API is compiled
// and exercised exactly as a reference user would call it.
//
//   dropin_demo <in.pgm> <sigma> <lobeSize> <outdir> [niter]
//
// Writes to <outdir>:
//   smooth_u8.pgm          FIRGaussianConvolver2<float>::smooth, u_char in/out (gaussTest)
//   <method>_f32.pfm       the six methods, u_char in, float out
//   fused_<method>.pfm     the same six from one convolve6() call
//   fir_f32.pfm            FIRFilter2 with computeGaussianCoefficients(sigma, lobeSize)
//                          lobes, float in (converted from u_char on upload) and out
//                          (filterTest)
// and prints one line of JSON with timings and the exception checks.
#include "TU/cu/FIRGaussianConvolver.h"
#include "TU/cu/algorithm.h"

#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <string>

namespace
{
// Testprog/filterTest/main.cu:17-47: lobeSize+1 coefficients, centre last,
// normalised to unit sum over the full kernel.
TU::Array<float>
computeGaussianCoefficients(float sigma, unsigned lobeSize)
{
    TU::Array<float> coeff(lobeSize + 1);
    for (unsigned i = 0; i <= lobeSize; ++i) {
        const float dx = float(lobeSize - i) / sigma;
        coeff[i] = float(std::exp(-0.5 * dx * dx));
    }
    float sum = coeff[lobeSize];
    for (unsigned i = 0; i < lobeSize; ++i) sum += (2.0f * coeff[i]);
    for (unsigned i = 0; i <= lobeSize; ++i) coeff[i] /= sum;
    return coeff;
}

template <class T>
void save(const std::string& path, const TU::cu::Array2<T>& a)
{
    TU::Image<T> img(a);
    std::ofstream out(path, std::ios::binary);
    img.save(out);
}

const char* const NAMES[6] = {"smooth", "diffH", "diffV", "diffHH", "diffHV", "diffVV"};
}   // namespace

int
main(int argc, char* argv[])
{
    using namespace TU;
    if (argc < 5) {
        std::cerr << "usage: dropin_demo in.pgm sigma lobeSize outdir [niter]" << std::endl;
        return 2;
    }
    const float       sigma    = float(std::atof(argv[2]));
    const unsigned    lobeSize = unsigned(std::atoi(argv[3]));
    const std::string dir      = argv[4];
    const int         niter    = argc > 5 ? std::atoi(argv[5]) : 100;

    try {
        Image<u_char> in;
        std::ifstream is(argv[1], std::ios::binary);
        in.restore(is);

        // gaussTest: u_char -> float intermediate -> u_char
        cu::FIRGaussianConvolver2<float> cudaFilter(sigma);
        cu::Array2<u_char>               in_d(in);
        cu::Array2<u_char>               out_d(in_d.nrow(), in_d.ncol());
        cudaFilter.smooth(in_d.cbegin(), in_d.cend(), out_d.begin());
        cu::check(tucu_synchronize());

        // timed like Profiler<cu::clock>: a synchronising clock per call
        double total = 0;
        for (int n = 0; n < niter; ++n) {
            const auto t0 = std::chrono::steady_clock::now();
            cudaFilter.smooth(in_d.cbegin(), in_d.cend(), out_d.begin());
            cu::check(tucu_synchronize());
            total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        save(dir + "/smooth_u8.pgm", out_d);

        // the six methods one by one, float out
        using Out = cu::Array2<float>;
        std::array<Out, 6> outs;
        for (auto& o : outs) o.resize(in_d.nrow(), in_d.ncol());
        cudaFilter.smooth(in_d.cbegin(), in_d.cend(), outs[0].begin());
        cudaFilter.diffH (in_d.cbegin(), in_d.cend(), outs[1].begin());
        cudaFilter.diffV (in_d.cbegin(), in_d.cend(), outs[2].begin());
        cudaFilter.diffHH(in_d.cbegin(), in_d.cend(), outs[3].begin());
        cudaFilter.diffHV(in_d.cbegin(), in_d.cend(), outs[4].begin());
        cudaFilter.diffVV(in_d.cbegin(), in_d.cend(), outs[5].begin());
        for (int k = 0; k < 6; ++k) save(dir + "/" + NAMES[k] + "_f32.pfm", outs[k]);

        // the pyramid step (Testprog/subsampleTest/main.cu:28-31 after a smooth)
        cu::Array2<float> half(in_d.nrow() / 2, in_d.ncol() / 2);
        cu::subsample(outs[0].cbegin(), outs[0].cend(), half.begin(), cu::identity1x1());
        save(dir + "/half_smooth.pfm", half);
        // the same step fused (extension): no full-size intermediate
        cu::Array2<float> half_fused(in_d.nrow() / 2, in_d.ncol() / 2);
        cudaFilter.smoothSubsample(in_d.cbegin(), in_d.cend(), half_fused.begin());
        save(dir + "/half_smooth_fused.pfm", half_fused);

        // the six from one fused launch
        std::array<Out, 6> fused;
        std::array<Out::iterator, 6> its;
        for (int k = 0; k < 6; ++k) { fused[k].resize(in_d.nrow(), in_d.ncol()); its[k] = fused[k].begin(); }
        cudaFilter.convolve6(in_d.cbegin(), in_d.cend(), its);
        for (int k = 0; k < 6; ++k) save(dir + "/fused_" + NAMES[k] + ".pfm", fused[k]);

        // filterTest: FIRFilter2 with explicit lobes, u_char converted to float on upload
        const Array<float>  coeff = computeGaussianCoefficients(sigma, lobeSize);
        cu::FIRFilter2<float> fir;
        fir.initialize(coeff, coeff);
        cu::Array2<float> inf_d(in);
        cu::Array2<float> outf_d(in.nrow(), in.ncol());
        fir.convolve(inf_d.cbegin(), inf_d.cend(), outf_d.begin());
        save(dir + "/fir_f32.pfm", outf_d);

        // a sub-image through cu::slice (TU/cu/iterator.h:588-609): rows
        // 10 .. nrow-11, columns 7 .. ncol-14 of the float upload, filtered
        // into the same window of another image; copy-initialised from a
        // float, as the reference's implicit constructor allows
        cu::FIRGaussianConvolver2<float> sub_filter = sigma;
        const int sr = 10, sc = 7, snr = int(inf_d.nrow()) - 21, snc = int(inf_d.ncol()) - 21;
        cu::Array2<float> sub_d(inf_d.nrow(), inf_d.ncol());
        const auto sub_in  = cu::slice(inf_d.cbegin(), sr, snr, sc, snc);
        const auto sub_out = cu::slice(sub_d.begin(), sr, snr, sc, snc);
        sub_filter.diffH(sub_in.begin(), sub_in.end(), sub_out.begin());
        const auto rows = cu::make_range(inf_d.cbegin() + 3, int(inf_d.nrow()) - 6);
        cu::Array2<float> rows_d(inf_d.nrow(), inf_d.ncol());
        sub_filter.smooth(rows.begin(), rows.end(), (rows_d.begin() + 3));
        save(dir + "/slice_diffH.pfm", sub_d);
        save(dir + "/range_smooth.pfm", rows_d);

        // one more subsample operator, on the float upload
        cu::Array2<float> g5(inf_d.nrow() / 2, inf_d.ncol() / 2);
        cu::subsample(inf_d.cbegin(), inf_d.cend(), g5.begin(), cu::gaussian5x5());
        save(dir + "/gaussian5x5_half.pfm", g5);

        // exceptions of the reference (FIRFilter.h:250-251, 344, 375)
        std::string too_large, unsupported;
        try { fir.initialize(Array<float>(18), Array<float>(3)); }
        catch (const std::runtime_error& e) { too_large = e.what(); }
        try {
            fir.initialize(Array<float>(6), Array<float>(3));
            fir.convolve(inf_d.cbegin(), inf_d.cend(), outf_d.begin());
        }
        catch (const std::runtime_error& e) { unsupported = e.what(); }
        cu::check(tucu_synchronize());

        std::cout << "{\"niter\": " << niter << ", \"smooth_ms\": " << total / niter
                  << ", \"nrow\": " << in.nrow() << ", \"ncol\": " << in.ncol()
                  << ", \"too_large\": \"" << too_large << "\", \"unsupported\": \"" << unsupported
                  << "\", \"offsetH\": " << fir.offsetH() << "}" << std::endl;
    }
    catch (const std::exception& err) {
        std::cerr << err.what() << std::endl;
        return 1;
    }
    return 0;
}
