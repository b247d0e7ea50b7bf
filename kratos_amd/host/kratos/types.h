// types.h -- scalar and 3-vector types, variable indices and loop ranges of
// the host API (the reference's types.h:19-26 / :133-136, dynamics
// indices.h:7-13 and utilities/mapping/loop.h:136-160).
//
// PRECISION selects the form like the reference's build flag: 2 (default)
// is the plain FP64 form, 1 the mixed form (FP32 float_t on the device's
// flux path, FP64 float2_t state).  PRECISION=0 has no counterpart here.
#pragma once

#include <cstddef>
#include <iterator>

#ifndef PRECISION
#define PRECISION 2
#endif

namespace type {
#if PRECISION == 2
typedef double float_t;
typedef double float2_t;
#elif PRECISION == 1
typedef float float_t;
typedef double float2_t;
#else
#error "PRECISION must be 1 (mixed form) or 2 (FP64 form)"
#endif
typedef int int_t;

template <class T>
struct v3_t {
    T x[3];
    v3_t() : x{T(0), T(0), T(0)} {}
    v3_t(T v) : x{v, v, v} {}
    v3_t(T a, T b, T c) : x{a, b, c} {}
    T &operator[](int a) { return x[a]; }
    const T &operator[](int a) const { return x[a]; }
    T prod() const { return x[0] * x[1] * x[2]; }
    bool operator==(const v3_t &o) const { return x[0] == o.x[0] && x[1] == o.x[1] && x[2] == o.x[2]; }
    bool operator!=(const v3_t &o) const { return !(*this == o); }
};
typedef v3_t<int_t> idx_t;
typedef v3_t<float_t> coord_t;
typedef v3_t<float2_t> coord2_t;
}  // namespace type

using type::coord2_t;
using type::coord_t;
using type::float2_t;
using type::idx_t;

namespace idx {
enum { rho = 0, ene = 1, mom = 2, bcc = 5 };
}  // namespace idx

namespace utils {
// Range over the box [lo, hi) with x fastest (the reference's utils::loop).
class loop_t {
public:
    struct iterator {
        type::idx_t i, lo, hi;
        bool end;
        const type::idx_t &operator*() const { return i; }
        iterator &operator++() {
            for (int a = 0; a < 3; ++a) {
                if (++i[a] < hi[a]) return *this;
                i[a] = lo[a];
            }
            end = true;
            return *this;
        }
        bool operator!=(const iterator &o) const { return end != o.end || (!end && i != o.i); }
    };
    loop_t(const type::idx_t &lo, const type::idx_t &hi) : lo_(lo), hi_(hi) {}
    iterator begin() const {
        bool empty = false;
        for (int a = 0; a < 3; ++a) empty |= hi_[a] <= lo_[a];
        return iterator{lo_, lo_, hi_, empty};
    }
    iterator end() const { return iterator{hi_, lo_, hi_, true}; }

private:
    type::idx_t lo_, hi_;
};
inline loop_t loop(const type::idx_t &hi) { return loop_t(type::idx_t(0), hi); }
inline loop_t loop(const type::idx_t &lo, const type::idx_t &hi) { return loop_t(lo, hi); }
}  // namespace utils
