/*!
  \file   Array++.h
  \brief  TU::cu::Array2<T>: 2-D device buffer with row iterators, the
          drop-in for the reference's cu::Array2 (TU/cu/Array++.h:188-197,
          TU/cu/allocator.h:58-97, TU/cu/iterator.h:334-611).

  Host-only C++ over the C ABI of libtucuda_amd.so (include/tucu.h): no HIP
  headers are needed to use it.  Storage follows cu::allocator: device
  memory zero-filled at allocation, rows padded to 256 bytes.
*/
#pragma once

#include "tucu.h"
#include "TU/cu/host.h"

#include <cstdlib>
#include <iostream>
#include <iterator>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace TU
{
namespace cu
{
//! Storage-only 16-bit float element types (converted to float on load).
struct half_t     { uint16_t bits; };
struct bfloat16_t { uint16_t bits; };

template <class T> struct dtype_of;
template <> struct dtype_of<unsigned char> { static constexpr int value = TUCU_U8; };
template <> struct dtype_of<float>         { static constexpr int value = TUCU_F32; };
template <> struct dtype_of<half_t>        { static constexpr int value = TUCU_F16; };
template <> struct dtype_of<bfloat16_t>    { static constexpr int value = TUCU_BF16; };

//! Turns a tucu_status into the exception the reference raises for the
//! same condition: std::runtime_error for lobe errors (FIRFilter.h:250-251,
//! 344, 375), std::bad_alloc for allocation (allocator.h:83-84).  Device
//! errors print and exit as gpuCheckLastError does (algorithm.h:85-94)
//! when TUCU_EXIT_ON_DEVICE_ERROR is defined, and throw otherwise.
inline void
check(int status, const char* file = __FILE__, int line = __LINE__)
{
    switch (status) {
      case TUCU_OK:
        return;
      case TUCU_ERR_BAD_ALLOC:
        throw std::bad_alloc();
      case TUCU_ERR_DEVICE:
#ifdef TUCU_EXIT_ON_DEVICE_ERROR
        std::cerr << "HIP Runtime Error at: " << file << ':' << line
                  << '[' << tucu_last_error() << ']' << std::endl;
        std::exit(EXIT_FAILURE);
#endif
        // fall through
      default:
        (void)file; (void)line;
        throw std::runtime_error(tucu_last_error());
    }
}

//! One row of a device image: pointer and length (what std::size(*row_it)
//! reports, as TU/cu/iterator.h:385-434's range does).
template <class T>
class row_range
{
  public:
    row_range(T* p, int n) : _p(p), _n(n) {}
    int  size()  const { return _n; }
    T*   data()  const { return _p; }
    T*   begin() const { return _p; }
    T*   end()   const { return _p + _n; }

  private:
    T*  _p;
    int _n;
};

//! Random-access iterator over the rows of a device image: row pointer,
//! stride in elements and row length (TU/cu/iterator.h:439-527).
template <class T>
class row_iterator
{
  public:
    using iterator_category = std::random_access_iterator_tag;
    using value_type        = row_range<T>;
    using difference_type   = std::ptrdiff_t;
    using reference         = row_range<T>;
    using pointer           = void;

    row_iterator() : _p(nullptr), _stride(0), _ncol(0) {}
    row_iterator(T* p, std::ptrdiff_t stride, int ncol) : _p(p), _stride(stride), _ncol(ncol) {}
    //! iterator -> const_iterator
    template <class S, class = std::enable_if_t<std::is_same<const S, T>::value && !std::is_same<S, T>::value>>
    row_iterator(const row_iterator<S>& it) : _p(it.get()), _stride(it.stride()), _ncol(it.ncol()) {}

    row_range<T>    operator *()                          const { return {_p, _ncol}; }
    row_range<T>    operator [](difference_type n)        const { return {_p + n * _stride, _ncol}; }
    row_iterator&   operator ++()                               { _p += _stride; return *this; }
    row_iterator&   operator --()                               { _p -= _stride; return *this; }
    row_iterator    operator ++(int)                            { auto t = *this; ++*this; return t; }
    row_iterator    operator --(int)                            { auto t = *this; --*this; return t; }
    row_iterator&   operator +=(difference_type n)              { _p += n * _stride; return *this; }
    row_iterator&   operator -=(difference_type n)              { _p -= n * _stride; return *this; }
    row_iterator    operator +(difference_type n)         const { return {_p + n * _stride, _stride, _ncol}; }
    row_iterator    operator -(difference_type n)         const { return {_p - n * _stride, _stride, _ncol}; }
    difference_type operator -(const row_iterator& it)    const { return _stride ? (_p - it._p) / _stride : 0; }
    bool            operator ==(const row_iterator& it)   const { return _p == it._p; }
    bool            operator !=(const row_iterator& it)   const { return _p != it._p; }
    bool            operator <(const row_iterator& it)    const { return _p < it._p; }

    T*              get()                                 const { return _p; }
    std::ptrdiff_t  stride()                              const { return _stride; }
    int             ncol()                                const { return _ncol; }

  private:
    T*             _p;
    std::ptrdiff_t _stride;
    int            _ncol;
};

//! A run of `size` rows starting at a row iterator: what cu::make_range and
//! cu::slice return (TU/cu/iterator.h:568-611), a view of device memory
//! that owns nothing.  begin() / end() are row iterators, so a range passes
//! to FIRFilter2::convolve and the Gaussian methods as an Array2 does.
template <class ITER>
class range
{
  public:
    using iterator       = ITER;
    using const_iterator = ITER;

    range(const ITER& begin, int size) : _begin(begin), _size(size) {}

    ITER begin()  const { return _begin; }
    ITER end()    const { return _begin + _size; }
    ITER cbegin() const { return begin(); }
    ITER cend()   const { return end(); }
    int  size()   const { return _size; }
    auto operator [](int i) const { return _begin[i]; }

  private:
    ITER _begin;
    int  _size;
};

//! make_range(iter, size) (TU/cu/iterator.h:568-573): `size` rows from iter.
template <class T> inline range<row_iterator<T>>
make_range(const row_iterator<T>& iter, int size)
{
    return {iter, size};
}

//! make_range(iter, size, stride, ncol): the same rows read with another
//! row stride (elements) and row length (TU/cu/iterator.h:575-579).
template <class T> inline range<row_iterator<T>>
make_range(const row_iterator<T>& iter, int size, std::ptrdiff_t stride, int ncol)
{
    return {row_iterator<T>(iter.get(), stride, ncol), size};
}

//! slice(iter, idx, size): rows idx .. idx + size - 1 (TU/cu/iterator.h:604-609).
template <class T> inline range<row_iterator<T>>
slice(const row_iterator<T>& iter, int idx, int size)
{
    return {iter + idx, size};
}

//! slice(iter, idx, size, jdx, jsize): the sub-image of rows idx .. idx +
//! size - 1 and columns jdx .. jdx + jsize - 1 (TU/cu/iterator.h:588-609;
//! the reference itself slices an output this way, FIRFilter.h:286-287).
template <class T> inline range<row_iterator<T>>
slice(const row_iterator<T>& iter, int idx, int size, int jdx, int jsize)
{
    const row_iterator<T> r = iter + idx;
    return {row_iterator<T>(r.get() + jdx, r.stride(), jsize), size};
}

//! The C descriptor of rows [ib, ie).
template <class T>
inline tucu_image
make_image(const row_iterator<T>& ib, const row_iterator<T>& ie)
{
    using E = std::remove_const_t<T>;
    return {const_cast<E*>(ib.get()), int64_t(ib.stride()), int32_t(ie - ib), int32_t(ib.ncol()),
            dtype_of<E>::value, 0};
}

//! Output descriptor: the row length is known from the iterator, the number
//! of rows is not (the reference takes the output extent on trust).
template <class T>
inline tucu_image
make_out_image(const row_iterator<T>& out)
{
    return {out.get(), int64_t(out.stride()), 0, int32_t(out.ncol()), dtype_of<T>::value, 0};
}

/************************************************************************
*  class Array2<T>                                                      *
************************************************************************/
//! 2-D array in device memory.
template <class T>
class Array2
{
  public:
    using value_type     = T;
    using iterator       = row_iterator<T>;
    using const_iterator = row_iterator<const T>;

    Array2() : _p(nullptr), _nrow(0), _ncol(0), _stride(0) {}
    Array2(size_t nrow, size_t ncol) : Array2() { resize(nrow, ncol); }
    //! Upload from a host 2-D array (nrow(), ncol(), operator[] rows) with
    //! element conversion, as cu::Array2<T>(Image<S>) does
    //! (Testprog/gaussTest/main.cu:91, Testprog/filterTest/main.cu:93).
    template <class HOST, class = decltype(std::declval<const HOST&>().nrow()),
              class = std::enable_if_t<!std::is_same<HOST, Array2>::value>>
    explicit Array2(const HOST& host) : Array2() { read(host); }
    Array2(const Array2& a) : Array2(a.nrow(), a.ncol())
    {
        check(tucu_memcpy2d(_p, _stride * sizeof(T), a._p, a._stride * sizeof(T),
                            _ncol * sizeof(T), _nrow, 2));
    }
    Array2(Array2&& a) noexcept : _p(a._p), _nrow(a._nrow), _ncol(a._ncol), _stride(a._stride)
    {
        a._p = nullptr; a._nrow = a._ncol = a._stride = 0;
    }
    Array2& operator =(Array2 a) noexcept
    {
        std::swap(_p, a._p); std::swap(_nrow, a._nrow);
        std::swap(_ncol, a._ncol); std::swap(_stride, a._stride);
        return *this;
    }
    ~Array2() { tucu_free(_p); }

    size_t  nrow()   const { return _nrow; }
    size_t  ncol()   const { return _ncol; }
    size_t  stride() const { return _stride; }       //!< elements between rows
    T*      data()         { return _p; }
    const T* data()  const { return _p; }

    //! Reallocates (zero-filled) when the shape changes.
    void resize(size_t nrow, size_t ncol)
    {
        if (nrow == _nrow && ncol == _ncol) return;
        tucu_free(_p);
        _p = nullptr; _nrow = _ncol = _stride = 0;
        const size_t stride = tucu_row_stride(ncol, sizeof(T));
        void* p = nullptr;
        check(tucu_malloc(&p, nrow * stride * sizeof(T)));
        _p = static_cast<T*>(p); _nrow = nrow; _ncol = ncol; _stride = stride;
    }

    iterator       begin()        { return {_p, std::ptrdiff_t(_stride), int(_ncol)}; }
    iterator       end()          { return begin() + std::ptrdiff_t(_nrow); }
    const_iterator begin()  const { return {_p, std::ptrdiff_t(_stride), int(_ncol)}; }
    const_iterator end()    const { return begin() + std::ptrdiff_t(_nrow); }
    const_iterator cbegin() const { return begin(); }
    const_iterator cend()   const { return end(); }

    //! Upload (host element type S converted to T on the host).
    template <class HOST> void read(const HOST& host)
    {
        resize(host.nrow(), host.ncol());
        std::vector<T> buf(_ncol * _nrow);
        for (size_t y = 0; y < _nrow; ++y)
            for (size_t x = 0; x < _ncol; ++x) buf[y * _ncol + x] = T(host[y][x]);
        check(tucu_memcpy2d(_p, _stride * sizeof(T), buf.data(), _ncol * sizeof(T),
                            _ncol * sizeof(T), _nrow, 0));
    }

    //! Download into a host 2-D array (resized to this shape), converting
    //! T to the host element type.
    template <class HOST> void write(HOST& host) const
    {
        std::vector<T> buf(_ncol * _nrow);
        check(tucu_memcpy2d(buf.data(), _ncol * sizeof(T), _p, _stride * sizeof(T),
                            _ncol * sizeof(T), _nrow, 1));
        host.resize(_nrow, _ncol);
        using E = std::remove_reference_t<decltype(host[0][0])>;
        for (size_t y = 0; y < _nrow; ++y)
            for (size_t x = 0; x < _ncol; ++x) host[y][x] = E(buf[y * _ncol + x]);
    }

  private:
    T*     _p;
    size_t _nrow, _ncol, _stride;
};

}   // namespace cu
}   // namespace TU
