/*!
  \file   host.h
  \brief  Minimal host containers standing in for TULibs' TU::Array,
          TU::Array2 and TU::Image (TULibs is a separate library the
          reference builds against; see Testprog/gaussTest/main.cu:85-87,
          108-113 for the uses mirrored here).

  Define TUCU_WITH_TULIBS before including the TU/cu headers to use the
  real TULibs types instead; the filters only need size()/data() of a lobe
  and nrow()/ncol()/operator[] of a host image.
*/
#pragma once

#include <cctype>
#include <cstddef>
#include <initializer_list>
#include <cstdint>
#include <cstring>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#ifndef TUCU_WITH_TULIBS
namespace TU
{
using u_char = unsigned char;

//! 1-D host array (the subset of TU::Array<T> the filters use).
template <class T>
class Array
{
  public:
    using value_type = T;

    explicit Array(size_t n = 0) : _v(n) {}
    Array(std::initializer_list<T> l) : _v(l) {}

    size_t      size()                   const { return _v.size(); }
    void        resize(size_t n)               { _v.resize(n); }
    T*          data()                         { return _v.data(); }
    const T*    data()                   const { return _v.data(); }
    T&          operator [](size_t i)          { return _v[i]; }
    const T&    operator [](size_t i)    const { return _v[i]; }
    auto        begin()                        { return _v.begin(); }
    auto        end()                          { return _v.end(); }
    auto        begin()                  const { return _v.begin(); }
    auto        end()                    const { return _v.end(); }

  private:
    std::vector<T> _v;
};

//! 2-D host array with contiguous rows.
template <class T>
class Array2
{
  public:
    using value_type = T;

    Array2(size_t nrow = 0, size_t ncol = 0) : _nrow(nrow), _ncol(ncol), _v(nrow * ncol) {}
    //! Converting copy from any 2-D source that can write itself into a host
    //! array (TU::cu::Array2 downloads this way).
    template <class SRC, class = decltype(std::declval<const SRC&>().write(std::declval<Array2&>()))>
    explicit Array2(const SRC& src) : Array2() { src.write(*this); }

    size_t      nrow()                   const { return _nrow; }
    size_t      ncol()                   const { return _ncol; }
    size_t      stride()                 const { return _ncol; }
    void        resize(size_t r, size_t c)     { _nrow = r; _ncol = c; _v.assign(r * c, T()); }
    T*          data()                         { return _v.data(); }
    const T*    data()                   const { return _v.data(); }
    T*          operator [](size_t y)          { return _v.data() + y * _ncol; }
    const T*    operator [](size_t y)    const { return _v.data() + y * _ncol; }

  private:
    size_t         _nrow, _ncol;
    std::vector<T> _v;
};

//! Grey-level host image with PNM I/O: 8-bit P5 / P6 and float Pf / PF
//! (PFM) in, P5 (byte images) or Pf (others) out.  Header comment lines
//! (TULibs' EPBM keys such as "# DataType: Char") are skipped, and a stream
//! may hold several images one after another, as the reference's
//! Testprog/ueshiba.epbm does: each restore() reads the next one.
template <class T>
class Image : public Array2<T>
{
  public:
    Image(size_t width = 0, size_t height = 0) : Array2<T>(height, width) {}
    template <class SRC, class = decltype(std::declval<const SRC&>().write(std::declval<Array2<T>&>()))>
    explicit Image(const SRC& src) : Array2<T>() { src.write(*this); }

    size_t width()  const { return this->ncol(); }
    size_t height() const { return this->nrow(); }

    //! Reads a binary PNM (P5 grey, P6 colour reduced to grey by the mean
    //! of the channels; Pf grey float, PF colour float reduced likewise,
    //! byte order from the sign of the scale: negative little-endian).
    std::istream& restore(std::istream& in)
    {
        std::string magic;
        in >> magic;
        auto skip = [&in] {
            for (int c; (c = in.peek()) != EOF; ) {
                if (c == '#') { std::string line; std::getline(in, line); }
                else if (std::isspace(c)) in.get();
                else break;
            }
        };
        size_t w, h;
        skip(); in >> w; skip(); in >> h;
        if (magic == "P5" || magic == "P6") {
            int maxval;
            skip(); in >> maxval;
            in.get();
            if (maxval > 255) throw std::runtime_error("Image::restore: 16-bit PNM is not supported");
            const size_t ch = magic == "P6" ? 3 : 1;
            std::vector<unsigned char> buf(w * h * ch);
            in.read(reinterpret_cast<char*>(buf.data()), std::streamsize(buf.size()));
            this->resize(h, w);
            for (size_t i = 0; i < w * h; ++i) {
                unsigned s = 0;
                for (size_t c = 0; c < ch; ++c) s += buf[i * ch + c];
                this->data()[i] = T(s / ch);
            }
        } else if (magic == "Pf" || magic == "PF") {
            double scale;
            skip(); in >> scale;
            in.get();
            const size_t ch = magic == "PF" ? 3 : 1;
            std::vector<uint32_t> buf(w * h * ch);
            in.read(reinterpret_cast<char*>(buf.data()), std::streamsize(buf.size() * sizeof(uint32_t)));
            const uint32_t one = 1;
            const bool host_little = *reinterpret_cast<const unsigned char*>(&one) == 1;
            if ((scale < 0) != host_little)             // file byte order differs from the host's
                for (uint32_t& v : buf)
                    v = (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
            this->resize(h, w);
            for (size_t y = 0; y < h; ++y)              // PFM stores the bottom row first
                for (size_t x = 0; x < w; ++x) {
                    float s = 0.f;
                    for (size_t c = 0; c < ch; ++c) {
                        float f;
                        std::memcpy(&f, &buf[(y * w + x) * ch + c], sizeof(f));
                        s += f;
                    }
                    (*this)[h - 1 - y][x] = T(s / float(ch));
                }
        } else {
            throw std::runtime_error("Image::restore: unsupported PNM type " + magic);
        }
        if (!in) throw std::runtime_error("Image::restore: truncated image");
        return in;
    }

    //! Writes P5 for byte images, Pf (little-endian grey PFM) otherwise.
    std::ostream& save(std::ostream& out) const
    {
        const size_t w = width(), h = height();
        if (std::is_same<T, unsigned char>::value) {
            out << "P5\n" << w << ' ' << h << "\n255\n";
            out.write(reinterpret_cast<const char*>(this->data()), std::streamsize(w * h));
        } else {
            out << "Pf\n" << w << ' ' << h << "\n-1.0\n";
            std::vector<float> buf(w * h);
            for (size_t y = 0; y < h; ++y)
                for (size_t x = 0; x < w; ++x) buf[y * w + x] = float((*this)[h - 1 - y][x]);
            out.write(reinterpret_cast<const char*>(buf.data()), std::streamsize(buf.size() * sizeof(float)));
        }
        return out;
    }
};
}   // namespace TU
#endif  // TUCU_WITH_TULIBS
