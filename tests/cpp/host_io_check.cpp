// Host image I/O of include/TU/cu/host.h (no GPU): reads every image of a
// PNM stream into Image<T>, prints "width height sum" per image, and writes
// the last one back out (P5 for u_char, Pf otherwise).
//   host_io_check u8|f32 in.pnm out.pnm
#include "TU/cu/host.h"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

template <class T>
int run(const char* in_path, const char* out_path)
{
    std::ifstream in(in_path, std::ios::binary);
    TU::Image<T> img;
    int n = 0;
    while (in.peek() != EOF) {
        img.restore(in);
        double sum = 0;
        for (size_t y = 0; y < img.height(); ++y)
            for (size_t x = 0; x < img.width(); ++x) sum += double(img[y][x]);
        std::printf("%zu %zu %.6f\n", img.width(), img.height(), sum);
        ++n;
        while (in.peek() == '\n' || in.peek() == ' ') in.get();
    }
    std::ofstream out(out_path, std::ios::binary);
    img.save(out);
    return n > 0 ? 0 : 1;
}

int main(int argc, char* argv[])
{
    if (argc != 4) { std::cerr << "usage: host_io_check u8|f32 in out" << std::endl; return 2; }
    try {
        return std::string(argv[1]) == "u8" ? run<unsigned char>(argv[2], argv[3]) : run<float>(argv[2], argv[3]);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
