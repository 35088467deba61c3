/* The AWCM compressible-Euler path behind the reference's C++ class names.
 *
 * doseongju/C_pp-library has no C interface: its drivers (Wavelet/
 * main_1d_euler.cpp, main_2d_euler(dmach).cpp, main_2d_euler(quad).cpp)
 * #include headers and call class methods.  This header provides those
 * classes with the same names, public members and method signatures, so such
 * a driver builds against this library with include/compat/ on the include
 * path (one forwarding header per reference header name).  The layouts are the
 * reference's: AoS VECTOR_3D/4D fields with ghost width 3, i fastest; one-byte
 * bool masks with ghost width 3; int levels without ghosts.  Fields therefore
 * cross to the GPU without repacking.
 *
 *   AWCM_1D / AWCM_2D::PrepareEuler  -> awcm_prepare_euler on the GPU; the
 *                                       field, masks and levels come back
 *   everything else                  -> host code in this file, with the
 *                                       reference's FP64 operation order
 *                                       (shared with the kernels through
 *                                       awcm/physics.h)
 *
 * A whole time loop on the GPU is awcm_run in awcm_capi.h; the per-point
 * stage loops of the reference drivers stay host-side here, as the reference
 * wrote them.
 *
 * Header-only C++17.  Include paths: include/ and c_pp-library_amd/include/;
 * link libawcm_hip.so.  Build with -ffp-contract=off for bit-identical
 * results.  Objects are not thread-safe, as in the reference; the solver
 * classes keep no scratch state, so one object may serve all threads.
 */
#pragma once

#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "awcm/physics.h"
#include "awcm_capi.h"

/* ---- MACROS.h ------------------------------------------------------------- */
#ifndef ABS
#define ABS(x) ((x) < 0 ? (-(x)) : (x))
#define SQR(x) ((x) * (x))
#define CUBE(x) ((x) * (x) * (x))
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define MIN(a, b) ((a) > (b) ? (b) : (a))
#define MAX3(a, b, c) (MAX(MAX(a, b), c))
#define MIN3(a, b, c) (MIN(MIN(a, b), c))
#define POSPART(x) ((x) > 0 ? (x) : (0))
#define NEGPART(x) ((x) < 0 ? (x) : (0))
#define CLAMP(v, lo, hi) (((v) < (lo)) ? (lo) : ((v) > (hi)) ? (hi) : (v))
#define PI (4. * atan(1.))
#endif
#define ITERATION_GRID_1D(field, width) for (int i = (field).i_start - (width); i <= (field).i_end + (width); i++)
#define ITERATION_GRID_2D(field, width)                                        \
    for (int j = (field).j_start - (width); j <= (field).j_end + (width); j++) \
        for (int i = (field).i_start - (width); i <= (field).i_end + (width); i++)
#define ITERATION_GRID_2D_IJ(field, width)                                     \
    for (int i = (field).i_start - (width); i <= (field).i_end + (width); i++) \
        for (int j = (field).j_start - (width); j <= (field).j_end + (width); j++)
#define ITERATION_GRID_INT_1D(grid, width) \
    for (int i = (grid).i_start - (width) * (grid).dx; i <= (grid).i_end + (width) * (grid).dx; i += (grid).dx)
#define ITERATION_GRID_INT_2D(sub, width)                                                                  \
    for (int j = (sub).j_start - (width) * (sub).dy; j <= (sub).j_end + (width) * (sub).dy; j += (sub).dy) \
        for (int i = (sub).i_start - (width) * (sub).dx; i <= (sub).i_end + (width) * (sub).dx; i += (sub).dx)
#ifndef SAFE_DELETE_ARRAY
#define SAFE_DELETE(p) \
    if ((p) != 0) {    \
        delete (p);    \
        (p) = 0;       \
    }
#define SAFE_DELETE_ARRAY(p) \
    if ((p) != 0) {          \
        delete[](p);         \
        (p) = 0;             \
    }
#endif

/* ---- small vectors (VECTOR_2D.h, VECTOR_3D.h, VECTOR_4D.h) ----------------
 * Every component is one FP64 operation, so sums of scaled vectors round as
 * in the reference.  VECTOR_3D * VECTOR_3D is component-wise and VECTOR_4D *
 * VECTOR_4D is the dot product, as the reference defines them. */
template <class T>
class VECTOR_2D {
public:
    union {
        struct { T i, j; };
        struct { T x, y; };
        T values[2];
    };
    VECTOR_2D() : i(T()), j(T()) {}
    VECTOR_2D(const T& x_, const T& y_) : x(x_), y(y_) {}
    VECTOR_2D(const VECTOR_2D<T>& v) : i(v.i), j(v.j) {}
    void operator=(const VECTOR_2D<T>& v) { i = v.i; j = v.j; }
    T& operator[](const int& k) { return values[k]; }
    const T& operator[](const int& k) const { return values[k]; }
};
template <class T>
std::ostream& operator<<(std::ostream& o, const VECTOR_2D<T>& v) { return o << v.i << " " << v.j; }

template <class T>
class VECTOR_3D {
public:
    union {
        struct { T i, j, k; };
        struct { T x, y, z; };
        T values[3];
    };
    VECTOR_3D() : i(0), j(0), k(0) {}
    VECTOR_3D(const T& a, const T& b, const T& c) : x(a), y(b), z(c) {}
    VECTOR_3D(const VECTOR_3D<T>& v) : i(v.i), j(v.j), k(v.k) {}
    void operator=(const VECTOR_3D<T>& v) { i = v.i; j = v.j; k = v.k; }
    T& operator[](const int& n) { return values[n]; }
    const T& operator[](const int& n) const { return values[n]; }
    void operator+=(const T& a) { i += a; j += a; k += a; }
    void operator-=(const T& a) { i -= a; j -= a; k -= a; }
    void operator*=(const T& a) { i *= a; j *= a; k *= a; }
    void operator/=(const T& a) { i /= a; j /= a; k /= a; }
    void operator+=(const VECTOR_3D<T>& v) { i += v.i; j += v.j; k += v.k; }
    void operator-=(const VECTOR_3D<T>& v) { i -= v.i; j -= v.j; k -= v.k; }
    void operator*=(const VECTOR_3D<T>& v) { i *= v.i; j *= v.j; k *= v.k; }
    void operator/=(const VECTOR_3D<T>& v) { i /= v.i; j /= v.j; k /= v.k; }
    VECTOR_3D<T> operator+(const VECTOR_3D<T>& v) const { return VECTOR_3D<T>(i + v.i, j + v.j, k + v.k); }
    VECTOR_3D<T> operator-(const VECTOR_3D<T>& v) const { return VECTOR_3D<T>(i - v.i, j - v.j, k - v.k); }
    VECTOR_3D<T> operator*(const VECTOR_3D<T>& v) const { return VECTOR_3D<T>(i * v.i, j * v.j, k * v.k); }
    VECTOR_3D<T> operator+(const T& c) const { return VECTOR_3D<T>(i + c, j + c, k + c); }
    VECTOR_3D<T> operator-(const T& c) const { return VECTOR_3D<T>(i - c, j - c, k - c); }
    VECTOR_3D<T> operator*(const T& c) const { return VECTOR_3D<T>(i * c, j * c, k * c); }
    VECTOR_3D<T> operator/(const T& c) const { return VECTOR_3D<T>(i / c, j / c, k / c); }
    double SquareMagnitude() const { return x * x + y * y + z * z; }
    double Magnitude() const { return std::sqrt(SquareMagnitude()); }
    void Normalize() {
        const T m = Magnitude();
        i /= m; j /= m; k /= m;
    }
};
template <class T>
std::ostream& operator<<(std::ostream& o, const VECTOR_3D<T>& v) { return o << v.i << " " << v.j << " " << v.k; }
template <class T>
T DotProduct(const VECTOR_3D<T>& a, const VECTOR_3D<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
VECTOR_3D<T> operator*(const T& c, const VECTOR_3D<T>& v) { return VECTOR_3D<T>(c * v[0], c * v[1], c * v[2]); }
template <class T>
VECTOR_3D<T> CrossProduct(const VECTOR_3D<T>& a, const VECTOR_3D<T>& b) {
    VECTOR_3D<T> r;
    r.x = a.y * b.z - a.z * b.y;
    r.y = -a.x * b.z + a.z * b.x;
    r.z = a.x * b.y - a.y * b.x;
    return r;
}

template <class T>
class VECTOR_4D {
public:
    T values[4];
    VECTOR_4D() { values[0] = values[1] = values[2] = values[3] = T(); }
    VECTOR_4D(const T& a, const T& b, const T& c, const T& d) { values[0] = a; values[1] = b; values[2] = c; values[3] = d; }
    VECTOR_4D(const VECTOR_4D<T>& v) { for (int n = 0; n < 4; n++) values[n] = v.values[n]; }
    void operator=(const VECTOR_4D<T>& v) { for (int n = 0; n < 4; n++) values[n] = v.values[n]; }
    T& operator[](const int& n) { return values[n]; }
    const T& operator[](const int& n) const { return values[n]; }
    void operator+=(const VECTOR_4D<T>& v) { for (int n = 0; n < 4; n++) values[n] += v[n]; }
    void operator-=(const VECTOR_4D<T>& v) { for (int n = 0; n < 4; n++) values[n] -= v[n]; }
    void operator*=(const VECTOR_4D<T>& v) { for (int n = 0; n < 4; n++) values[n] *= v[n]; }
    void operator*=(const T& c) { for (int n = 0; n < 4; n++) values[n] *= c; }
    void operator/=(const VECTOR_4D<T>& v) { for (int n = 0; n < 4; n++) values[n] /= v[n]; }
    VECTOR_4D<T> operator+(const T& c) const { return VECTOR_4D<T>(values[0] + c, values[1] + c, values[2] + c, values[3] + c); }
    VECTOR_4D<T> operator-(const T& c) const { return VECTOR_4D<T>(values[0] - c, values[1] - c, values[2] - c, values[3] - c); }
    VECTOR_4D<T> operator*(const T& c) const { return VECTOR_4D<T>(values[0] * c, values[1] * c, values[2] * c, values[3] * c); }
    VECTOR_4D<T> operator/(const T& c) const { return VECTOR_4D<T>(values[0] / c, values[1] / c, values[2] / c, values[3] / c); }
    VECTOR_4D<T> operator+(const VECTOR_4D<T>& v) const {
        return VECTOR_4D<T>(values[0] + v[0], values[1] + v[1], values[2] + v[2], values[3] + v[3]);
    }
    VECTOR_4D<T> operator-(const VECTOR_4D<T>& v) const {
        return VECTOR_4D<T>(values[0] - v[0], values[1] - v[1], values[2] - v[2], values[3] - v[3]);
    }
    T operator*(const VECTOR_4D<T>& v) const {
        return T(values[0] * v[0] + values[1] * v[1] + values[2] * v[2] + values[3] * v[3]);
    }
    double SquareMagnitude() const { return (*this) * (*this); }
    double Magnitude() const { return std::sqrt(SquareMagnitude()); }
};
template <class T>
std::ostream& operator<<(std::ostream& o, const VECTOR_4D<T>& v) { return o << v[0] << " " << v[1] << " " << v[2] << " " << v[3]; }
template <class T>
T DotProduct(const VECTOR_4D<T>& a, const VECTOR_4D<T>& b) { return a * b; }
template <class T>
VECTOR_4D<T> operator*(const T& c, const VECTOR_4D<T>& v) { return VECTOR_4D<T>(c * v[0], c * v[1], c * v[2], c * v[3]); }

static_assert(sizeof(VECTOR_3D<double>) == 24 && sizeof(VECTOR_4D<double>) == 32, "AoS state layout");
static_assert(sizeof(bool) == 1 && sizeof(int) == 4, "mask and level layout");

/* ---- owning arrays (ARRAY.h, ARRAY_1D.h, ARRAY_2D.h) -----------------------
 * Public storage under the reference's member names (values_, val).  Copies
 * are deep; the reference's ARRAY_1D copy constructor allocates without
 * copying, which no driver relies on. */
template <class TT>
class ARRAY {
public:
    int num_elements_ = 0;
    TT* values_ = nullptr;
    ARRAY() {}
    explicit ARRAY(const int& n) { Initialize(n); }
    ARRAY(const int& n, const TT& v) { Initialize(n, v); }
    ARRAY(const ARRAY<TT>& a) { Initialize(a); }
    ~ARRAY() { SAFE_DELETE_ARRAY(values_); }
    void Initialize(const int& n) {
        SAFE_DELETE_ARRAY(values_);
        num_elements_ = n;
        values_ = new TT[n];
    }
    void Initialize(const int& n, const TT& v) {
        Initialize(n);
        AssignAllValues(v);
    }
    void Initialize(const ARRAY<TT>& a) {
        if (this == &a) return;
        Initialize(a.num_elements_);
        for (int k = 0; k < num_elements_; k++) values_[k] = a.values_[k];
    }
    void AssignAllValues(const TT& v) { for (int k = 0; k < num_elements_; k++) values_[k] = v; }
    void AssignValues(const int& s, const int& e, const TT& v) { for (int k = s; k <= e; k++) values_[k] = v; }
    void CopyFrom(const ARRAY<TT>& a) { for (int k = 0; k < num_elements_; k++) values_[k] = a.values_[k]; }
    TT& operator[](const int& k) { return values_[k]; }
    const TT& operator[](const int& k) const { return values_[k]; }
    ARRAY<TT>& operator=(const ARRAY<TT>& a) { Initialize(a); return *this; }
    void Print(const char* filename) const {
        std::ofstream f(filename);
        for (int k = 0; k < num_elements_; k++) f << values_[k] << std::endl;
    }
};

template <class T>
class ARRAY_1D {
public:
    int i_start = 0, i_end = 0, i_res = 0;
    T* val = nullptr;
    ARRAY_1D() {}
    ARRAY_1D(const int& i_start_, const int& i_res_) { Initialize(i_start_, i_res_, true); }
    ARRAY_1D(const ARRAY_1D<T>& a) { *this = a; }
    ~ARRAY_1D() { SAFE_DELETE_ARRAY(val); }
    void Initialize(const int& i_start_, const int& i_res_, const bool& initialize = false) {
        i_start = i_start_;
        i_res = i_res_;
        i_end = i_start + i_res - 1;
        SAFE_DELETE_ARRAY(val);
        val = new T[i_res];
        if (initialize)
            for (int k = 0; k < i_res; k++) val[k] = T();
    }
    // named by the drivers (main_2d_euler(dmach).cpp:86) and by
    // FIELD_UNIFORM_1D.h:67, though the reference class does not define it
    void AssignAllValue(const T& v) { for (int k = 0; k < i_res; k++) val[k] = v; }
    T& operator()(const int& i) {
        assert(i >= i_start && i <= i_end);
        return val[i - i_start];
    }
    const T& operator()(const int& i) const {
        assert(i >= i_start && i <= i_end);
        return val[i - i_start];
    }
    ARRAY_1D<T>& operator=(const ARRAY_1D<T>& a) {
        if (this == &a) return *this;
        Initialize(a.i_start, a.i_res);
        for (int k = 0; k < i_res; k++) val[k] = a.val[k];
        return *this;
    }
};

template <class T>
class ARRAY_2D {
public:
    int i_start = 0, i_end = 0, j_start = 0, j_end = 0, i_res = 0, j_res = 0, ij_res = 0;
    T* val = nullptr;
    ARRAY_2D() {}
    ARRAY_2D(const int& i_start_, const int& j_start_, const int& i_res_, const int& j_res_,
             const bool& initialize = false) {
        Initialize(i_start_, j_start_, i_res_, j_res_, initialize);
    }
    ARRAY_2D(const ARRAY_2D<T>& a) { *this = a; }
    ~ARRAY_2D() { SAFE_DELETE_ARRAY(val); }
    void Initialize(const int& i_start_, const int& j_start_, const int& i_res_, const int& j_res_,
                    const bool& initialize = false) {
        i_start = i_start_; i_res = i_res_; i_end = i_start + i_res - 1;
        j_start = j_start_; j_res = j_res_; j_end = j_start + j_res - 1;
        ij_res = i_res * j_res;
        SAFE_DELETE_ARRAY(val);
        val = new T[ij_res];
        if (initialize)
            for (int k = 0; k < ij_res; k++) val[k] = T();
    }
    void Initialize(const ARRAY_2D<T>& a, const bool& initialize = false) {
        Initialize(a.i_start, a.j_start, a.i_res, a.j_res, initialize);
    }
    void AssignAllValues(const T& v) { for (int k = 0; k < ij_res; k++) val[k] = v; }
    // linear index of ARRAY_2D.h:112, i fastest
    T& operator()(const int& i, const int& j) {
        assert(i >= i_start && i <= i_end && j >= j_start && j <= j_end);
        return val[(i - i_start) + (j - j_start) * i_res];
    }
    const T& operator()(const int& i, const int& j) const {
        assert(i >= i_start && i <= i_end && j >= j_start && j <= j_end);
        return val[(i - i_start) + (j - j_start) * i_res];
    }
    ARRAY_2D<T>& operator=(const ARRAY_2D<T>& a) {
        if (this == &a) return *this;
        Initialize(a.i_start, a.j_start, a.i_res, a.j_res);
        for (int k = 0; k < ij_res; k++) val[k] = a.val[k];
        return *this;
    }
};

/* ---- uniform grids (GRID_UNIFORM_1D.h, GRID_UNIFORM_2D.h) ----------------- */
class GRID_UNIFORM_1D {
public:
    double x_min = 0, x_max = 0;
    double dx = 0, one_over_dx = 0;
    int i_start = 0, i_end = 0, i_res = 0;
    GRID_UNIFORM_1D() {}
    GRID_UNIFORM_1D(const int& i_start_, const int& i_res_, const double& x_min_, const double& x_max_) {
        Initialize(i_start_, i_res_, x_min_, x_max_);
    }
    void Initialize(const int& i_start_, const int& i_res_, const double& x_min_, const double& x_max_) {
        i_start = i_start_; i_res = i_res_; i_end = i_start + i_res - 1;
        x_min = x_min_; x_max = x_max_;
        dx = (x_max - x_min) / (i_res - 1);
        one_over_dx = 1 / dx;
    }
    void Initialize(const GRID_UNIFORM_1D& g) { Initialize(g.i_start, g.i_res, g.x_min, g.x_max); }
    double operator()(const int& i) const { return x_min + dx * (i - i_start); }
    int ClampIndexI(const int& i) const { return i < i_start ? i_start : (i > i_end ? i_end : i); }
    bool IsInside(const int& i) const { return i_start <= i && i <= i_end; }
    void PrintGrid(const char* filename) const {
        std::ofstream f(filename);
        for (int i = i_start; i <= i_end; i++) f << x_min + dx * (i - i_start) << " ";
    }
};

class GRID_UNIFORM_2D {
public:
    int i_start = 0, i_res = 0, i_end = 0, j_start = 0, j_res = 0, j_end = 0;
    double x_min = 0, x_max = 0, y_min = 0, y_max = 0;
    double dx = 0, one_over_dx = 0, dy = 0, one_over_dy = 0;
    GRID_UNIFORM_2D() {}
    GRID_UNIFORM_2D(const int& i_start_, const int& j_start_, const int& i_res_, const int& j_res_, const double& x_min_,
                    const double& y_min_, const double& x_max_, const double& y_max_) {
        Initialize(i_start_, j_start_, i_res_, j_res_, x_min_, y_min_, x_max_, y_max_);
    }
    void Initialize(const int& i_start_, const int& j_start_, const int& i_res_, const int& j_res_, const double& x_min_,
                    const double& y_min_, const double& x_max_, const double& y_max_) {
        i_start = i_start_; i_res = i_res_; i_end = i_start + i_res - 1;
        j_start = j_start_; j_res = j_res_; j_end = j_start + j_res - 1;
        x_min = x_min_; x_max = x_max_; y_min = y_min_; y_max = y_max_;
        dx = (x_max - x_min) / (i_res - 1); one_over_dx = 1 / dx;
        dy = (y_max - y_min) / (j_res - 1); one_over_dy = 1 / dy;
    }
    void Initialize(const GRID_UNIFORM_2D& g) {
        Initialize(g.i_start, g.j_start, g.i_res, g.j_res, g.x_min, g.y_min, g.x_max, g.y_max);
    }
    VECTOR_2D<double> operator()(const int& i, const int& j) const {
        return VECTOR_2D<double>(x_min + (i - i_start) * dx, y_min + (j - j_start) * dy);
    }
    void PrintGridX(const char* filename) const {
        std::ofstream f(filename);
        for (int i = i_start; i <= i_end; i++) f << x_min + dx * (i - i_start) << " ";
    }
    void PrintGridY(const char* filename) const {
        std::ofstream f(filename);
        for (int j = j_start; j <= j_end; j++) f << y_min + dy * (j - j_start) << " ";
    }
};

/* ---- fields (FIELD_UNIFORM_1D.h, FIELD_UNIFORM_2D.h) ----------------------
 * Print(bool) numbers its files with one counter per field type, "%s%4.4d"
 * in 1-D and "%s_%4.4d" in 2-D, as the reference names them. */
template <class T>
class FIELD_UNIFORM_1D {
public:
    ARRAY_1D<T> arr;
    GRID_UNIFORM_1D base_grid;
    int i_start = 0, i_res = 0, i_end = 0;
    int i_start_g = 0, i_res_g = 0, i_end_g = 0;
    int ghost_width = 0;
    const char* name = nullptr;
    FIELD_UNIFORM_1D() {}
    FIELD_UNIFORM_1D(const int& i_start_, const int& i_res_, const double& x_min, const double& x_max,
                     const int& ghost_width_ = 0) {
        Initialize(i_start_, i_res_, x_min, x_max, ghost_width_);
    }
    FIELD_UNIFORM_1D(const GRID_UNIFORM_1D& g, const int& ghost_width_ = 0) { Initialize(g, ghost_width_); }
    // copies the interior only, with a new ghost width (FIELD_UNIFORM_1D.h:30-38)
    FIELD_UNIFORM_1D(const FIELD_UNIFORM_1D<T>& f, const int& ghost_width_ = 0) {
        Initialize(f.base_grid, ghost_width_);
        for (int i = f.base_grid.i_start; i <= f.base_grid.i_end; i++) arr(i) = f(i);
    }
    FIELD_UNIFORM_1D<T>& operator=(const FIELD_UNIFORM_1D<T>& f) = default;
    void Initialize(const int& i_start_, const int& i_res_, const double& x_min, const double& x_max,
                    const int& ghost_width_ = 0) {
        ghost_width = ghost_width_;
        i_start = i_start_; i_res = i_res_; i_end = i_start + i_res - 1;
        i_start_g = i_start - ghost_width; i_res_g = i_res + 2 * ghost_width; i_end_g = i_end + ghost_width;
        arr.Initialize(i_start_g, i_res_g, true);
        base_grid.Initialize(i_start, i_res, x_min, x_max);
    }
    void Initialize(const GRID_UNIFORM_1D& g, const int& ghost_width_ = 0) {
        Initialize(g.i_start, g.i_res, g.x_min, g.x_max, ghost_width_);
    }
    void AssignAllValues(const T& v) { arr.AssignAllValue(v); }
    void TagName(const char* name_) { name = name_; }
    T& operator()(const int& i) { return arr(i); }
    const T& operator()(const int& i) const { return arr(i); }
    void Print(const char* filename, const bool ghost_cell = false) const {
        std::ofstream f(filename);
        const int a = ghost_cell ? i_start_g : i_start, b = ghost_cell ? i_end_g : i_end;
        for (int i = a; i <= b; i++) f << arr(i) << " ";
    }
    void Print(const bool& ghost_cell = false) const {
        static int cnt = 0;
        char filename[64];
        if (name) std::snprintf(filename, sizeof filename, "%s%4.4d", name, cnt);
        else std::snprintf(filename, sizeof filename, "%4.4d", cnt);
        Print(filename, ghost_cell);
        cnt++;
    }
};

template <class T>
class FIELD_UNIFORM_2D {
public:
    ARRAY_2D<T> arr;
    GRID_UNIFORM_2D grid;
    int i_start = 0, i_res = 0, i_end = 0, j_start = 0, j_res = 0, j_end = 0, ij_res = 0;
    int i_start_g = 0, i_res_g = 0, i_end_g = 0, j_start_g = 0, j_res_g = 0, j_end_g = 0, ij_res_g = 0;
    int ghost_width = 0;
    const char* name = nullptr;
    FIELD_UNIFORM_2D() {}
    FIELD_UNIFORM_2D(const int& i_start_, const int& j_start_, const int& i_res_, const int& j_res_, const double& x_min_,
                     const double& y_min_, const double& x_max_, const double& y_max_, const int& ghost_width_ = 0) {
        Initialize(i_start_, j_start_, i_res_, j_res_, x_min_, y_min_, x_max_, y_max_, ghost_width_);
    }
    FIELD_UNIFORM_2D(const GRID_UNIFORM_2D& g, const int& ghost_width_ = 0) { Initialize(g, ghost_width_); }
    void Initialize(const int& i_start_, const int& j_start_, const int& i_res_, const int& j_res_, const double& x_min_,
                    const double& y_min_, const double& x_max_, const double& y_max_, const int& ghost_width_ = 0) {
        ghost_width = ghost_width_;
        i_start = i_start_; i_res = i_res_; i_end = i_start + i_res - 1;
        j_start = j_start_; j_res = j_res_; j_end = j_start + j_res - 1;
        ij_res = i_res * j_res;
        i_start_g = i_start - ghost_width; i_res_g = i_res + 2 * ghost_width; i_end_g = i_end + ghost_width;
        j_start_g = j_start - ghost_width; j_res_g = j_res + 2 * ghost_width; j_end_g = j_end + ghost_width;
        ij_res_g = i_res_g * j_res_g;
        grid.Initialize(i_start, j_start, i_res, j_res, x_min_, y_min_, x_max_, y_max_);
        arr.Initialize(i_start_g, j_start_g, i_res_g, j_res_g, true);
    }
    void Initialize(const GRID_UNIFORM_2D& g, const int& ghost_width_ = 0) {
        Initialize(g.i_start, g.j_start, g.i_res, g.j_res, g.x_min, g.y_min, g.x_max, g.y_max, ghost_width_);
    }
    void AssignAllValues(const T& v) { arr.AssignAllValues(v); }
    void TagName(const char* name_) { name = name_; }
    T& operator()(const int& i, const int& j) { return arr(i, j); }
    const T& operator()(const int& i, const int& j) const { return arr(i, j); }
    void Print(const char* filename, const bool print_ghost = false) const {
        std::ofstream f(filename);
        f.precision(15);
        const int ia = print_ghost ? i_start_g : i_start, ib = print_ghost ? i_end_g : i_end;
        const int ja = print_ghost ? j_start_g : j_start, jb = print_ghost ? j_end_g : j_end;
        for (int j = ja; j <= jb; j++) {
            for (int i = ia; i <= ib; i++) f << (*this)(i, j) << " ";
            f << std::endl;
        }
    }
    void Print(const bool& print_ghost = false) const {
        static int cnt = 0;
        char filename[64];
        if (name) std::snprintf(filename, sizeof filename, "%s_%4.4d", name, cnt);
        else std::snprintf(filename, sizeof filename, "%4.4d", cnt);
        Print(filename, print_ghost);
        cnt++;
    }
};

/* ---- MRA integer grids (Wavelet/MRA_GRID_UNIFORM_1D.h, _2D.h) -------------
 * The per-level index ranges of the dyadic hierarchy.  The integer formulas
 * are the reference's, since they fix which points each level owns. */
class GRID_INT_1D {
public:
    int i_start = 0, i_end = 0, i_res = 0;
    int dx = 0, dx2 = 0;
    GRID_INT_1D() {}
    GRID_INT_1D(const int& s, const int& e, const int& r) { Initialize(s, e, r); }
    void Initialize(const int& s, const int& e, const int& r) {
        i_start = s; i_end = e; i_res = r;
        dx = (i_end - i_start) / (i_res - 1);
        dx2 = dx * 2;
    }
    bool IsInside(const int& i, const int& width = 0) const { return !(i < i_start - width || i_end + width < i); }
};
inline std::ostream& operator<<(std::ostream& o, const GRID_INT_1D& g) {
    return o << g.i_start << " " << g.i_end << " " << g.i_res << " " << g.dx;
}

class GRID_INT_2D {
public:
    int i_start = 0, i_end = 0, i_res = 0;
    int j_start = 0, j_end = 0, j_res = 0;
    int dx = 0, dx2 = 0;
    int dy = 0, dy2 = 0;
    GRID_INT_2D() {}
    GRID_INT_2D(const int& is, const int& js, const int& ie, const int& je, const int& ir, const int& jr) {
        Initialize(is, js, ie, je, ir, jr);
    }
    void Initialize(const int& is, const int& js, const int& ie, const int& je, const int& ir, const int& jr) {
        i_start = is; j_start = js; i_end = ie; j_end = je; i_res = ir; j_res = jr;
        dx = (i_end - i_start) / (i_res - 1);
        dy = (j_end - j_start) / (j_res - 1);
        dx2 = dx * 2;
        dy2 = dy * 2;
    }
    bool IsInside(const int& i, const int& j, const int& width = 0) const {
        return !(i < i_start - width || i_end + width < i || j < j_start - width || j_end + width < j);
    }
};
inline std::ostream& operator<<(std::ostream& o, const GRID_INT_2D& g) {
    return o << g.i_start << " " << g.j_start << " " << g.i_end << " " << g.j_end << " " << g.i_res << " " << g.j_res
             << " " << g.dx << " " << g.dy;
}

class MRA_GRID_INT_1D {
public:
    int num_levels = 0, Nx_coarse = 0, Nx = 0;
    ARRAY<int> dx;
    ARRAY<GRID_INT_1D> grid, grid_x;
    GRID_INT_1D grid_finest;
    void Initialize(const int& num_levels_, const int& Nx_coarse_) {
        num_levels = num_levels_;
        Nx_coarse = Nx_coarse_;
        Nx = (Nx_coarse - 1) * (1 << (num_levels - 1)) + 1;
        dx.Initialize(num_levels);
        grid.Initialize(num_levels);
        grid_x.Initialize(num_levels - 1);
        for (int lev = num_levels - 1; lev >= 0; lev--) {
            dx[lev] = 1 << (num_levels - lev - 1);
            if (lev == num_levels - 1) {
                grid[lev].Initialize(0, Nx - 1, Nx);
            } else {
                grid[lev].Initialize(0, Nx - 1, (Nx_coarse - 1) * (1 << lev) + 1);
                grid_x[lev].Initialize(dx[lev + 1], (Nx - 1) - dx[lev + 1], (Nx - 1) >> (num_levels - 1 - lev));
            }
        }
        grid_finest = grid[num_levels - 1];
    }
};

class MRA_GRID_INT_2D {
public:
    int num_levels = 0, Nx_coarse = 0, Ny_coarse = 0, Nx = 0, Ny = 0;
    ARRAY<int> mra;
    ARRAY<GRID_INT_2D> grid;
    ARRAY_2D<GRID_INT_2D> subgrid;  // (lev, mu), lev < num_levels-1, mu = 1..3
    GRID_INT_2D grid_finest;
    void Initialize(const int& num_levels_, const int& Nx_coarse_, const int& Ny_coarse_) {
        num_levels = num_levels_;
        Nx_coarse = Nx_coarse_;
        Ny_coarse = Ny_coarse_;
        Nx = (Nx_coarse - 1) * (1 << (num_levels - 1)) + 1;
        Ny = (Ny_coarse - 1) * (1 << (num_levels - 1)) + 1;
        mra.Initialize(num_levels);
        grid.Initialize(num_levels);
        subgrid.Initialize(0, 1, num_levels - 1, 3);
        for (int lev = num_levels - 1; lev >= 0; lev--) {
            const int m = 1 << (num_levels - lev - 1);
            mra[lev] = m;
            const int rx = (Nx_coarse - 1) * (1 << lev) + 1, ry = (Ny_coarse - 1) * (1 << lev) + 1;
            if (lev == num_levels - 1) {
                grid[lev].Initialize(0, 0, Nx - 1, Ny - 1, rx, ry);
                continue;
            }
            grid[lev].Initialize(m, m, (Nx - 1) - m, (Ny - 1) - m, rx - 2, ry - 2);
            const int h = m / 2;  // m >= 2 below the finest level
            subgrid(lev, 1).Initialize(h, m, Nx - h - 1, Ny - m - 1, Nx / m, Ny / m - 1);
            subgrid(lev, 2).Initialize(m, h, Nx - m - 1, Ny - h - 1, Nx / m - 1, Ny / m);
            subgrid(lev, 3).Initialize(h, h, Nx - h - 1, Ny - h - 1, Nx / m, Ny / m);
        }
        grid_finest = grid[num_levels - 1];
    }
};

/* ---- equation of state, physical flux, Roe solver (IDEAL_GAS_EQUATIONS.h,
 *      EULER_EQUATION.h, ROESOLVER.h) ---------------------------------------- */
class IDEAL_GAS_1D {
public:
    const double gamma;
    IDEAL_GAS_1D(const double& gamma_ = 1.4) : gamma(gamma_) {}
    double E(const double& rho, const double& u, const double& p) const { return awcm::eos_E1(gamma, rho, u, p); }
    double e(const double& rho, const double& u, const double& p) const { return E(rho, u, p) / rho - 0.5 * SQR(u); }
    double H(const double& rho, const double& u, const double& p) const { return (E(rho, u, p) + p) / rho; }
    double c(const double& rho, const double& p) const { return std::sqrt(gamma * p / rho); }
    double E(const VECTOR_3D<double>& q) const { return q[2]; }
    double e(const VECTOR_3D<double>& q) const { return q[2] / q[0] - 0.5 * SQR(q[1] / q[0]); }
    double p(const VECTOR_3D<double>& q) const { return awcm::eos_p1(gamma, q.values); }
    double H(const VECTOR_3D<double>& q) const { return awcm::eos_H1(gamma, q.values); }
    double c(const VECTOR_3D<double>& q) const { return awcm::eos_c1(gamma, q.values); }
};

class IDEAL_GAS_2D {
public:
    const double gamma;
    IDEAL_GAS_2D(const double& gamma_ = 1.4) : gamma(gamma_) {}
    double E(const double& rho, const double& u, const double& v, const double& p) const {
        return awcm::eos_E2(gamma, rho, u, v, p);
    }
    double e(const double& rho, const double& u, const double& v, const double& p) const {
        return E(rho, u, v, p) / rho - 0.5 * (u * u + v * v);
    }
    double H(const double& rho, const double& u, const double& v, const double& p) const {
        return (E(rho, u, v, p) + p) / rho;
    }
    double c(const double& rho, const double& p) const { return std::sqrt(gamma * p / rho); }
    double E(const VECTOR_4D<double>& q) const { return q[3]; }
    double e(const VECTOR_4D<double>& q) const { return q[3] / q[0] - 0.5 * (SQR(q[1]) + SQR(q[2])) / SQR(q[0]); }
    double p(const VECTOR_4D<double>& q) const { return awcm::eos_p2(gamma, q.values); }
    double H(const VECTOR_4D<double>& q) const { return awcm::eos_H2(gamma, q.values); }
    double c(const VECTOR_4D<double>& q) const { return awcm::eos_c2(gamma, q.values); }
};

class EULER_EQUATION_1D {
public:
    const double gamma;
    const IDEAL_GAS_1D EOS;
    EULER_EQUATION_1D(const double& gamma_ = 1.4) : gamma(gamma_), EOS(gamma_) {}
    void ComputeFu(const VECTOR_3D<double>& Q, VECTOR_3D<double>& fu) const { awcm::flux_f1(gamma, Q.values, fu.values); }
    void ConservedToPrimitive(const VECTOR_3D<double>& Q, double& rho, double& u, double& p) const {
        rho = Q[0];
        u = Q[1] / Q[0];
        p = EOS.p(Q);
    }
    void PrimitiveToConserved(const double& rho, const double& u, const double& p, VECTOR_3D<double>& Q) const {
        Q[0] = rho;
        Q[1] = rho * u;
        Q[2] = awcm::eos_E1(gamma, rho, u, p);
    }
    double Spectral_Radius(const VECTOR_3D<double>& q) const { return awcm::spec_rad1(gamma, q.values); }
    double MaxEigenvalue(const VECTOR_3D<double>& q) const { return q[1] / q[0] + EOS.c(q); }
    double MinEigenvalue(const VECTOR_3D<double>& q) const { return q[1] / q[0] - EOS.c(q); }
};

class EULER_EQUATION_2D {
public:
    const double gamma;
    const IDEAL_GAS_2D EOS;
    EULER_EQUATION_2D(const double& gamma_ = 1.4) : gamma(gamma_), EOS(gamma_) {}
    void ComputeFu(const VECTOR_4D<double>& Q, VECTOR_4D<double>& fu) const { awcm::flux_f2<0>(gamma, Q.values, fu.values); }
    void ComputeGu(const VECTOR_4D<double>& Q, VECTOR_4D<double>& gu) const { awcm::flux_f2<1>(gamma, Q.values, gu.values); }
    void ConservedToPrimitive(const VECTOR_4D<double>& Q, double& rho, double& u, double& v, double& p) const {
        rho = Q[0];
        u = Q[1] / Q[0];
        v = Q[2] / Q[0];
        p = EOS.p(Q);
    }
    void PrimitiveToConserved(const double& rho, const double& u, const double& v, const double& p,
                              VECTOR_4D<double>& Q) const {
        Q[0] = rho;
        Q[1] = rho * u;
        Q[2] = rho * v;
        Q[3] = awcm::eos_E2(gamma, rho, u, v, p);
    }
    double Spectral_RadiusX(const VECTOR_4D<double>& q) const { return awcm::spec_rad2x(gamma, q.values); }
    double Spectral_RadiusY(const VECTOR_4D<double>& q) const { return awcm::spec_rad2y(gamma, q.values); }
    double MaxEigenvalueX(const VECTOR_4D<double>& q) const { return q[1] / q[0] + EOS.c(q); }
    double MinEigenvalueX(const VECTOR_4D<double>& q) const { return q[1] / q[0] - EOS.c(q); }
    double MaxEigenvalueY(const VECTOR_4D<double>& q) const { return q[2] / q[0] + EOS.c(q); }
    double MinEigenvalueY(const VECTOR_4D<double>& q) const { return q[2] / q[0] - EOS.c(q); }
};

class ROE_SOLVER_1D {
public:
    static void RoeAverage(const VECTOR_3D<double>& Q_l, const VECTOR_3D<double>& Q_r, double& u, double& H, double& c,
                           const double& gamma = 1.4) {
        const awcm::Roe1 a = awcm::roe_avg1(gamma, Q_l.values, Q_r.values);
        u = a.u; H = a.H; c = a.c;
    }
    static void Inversion(const double& u, const double& H, const double& c, const double& b1, const double& b2,
                          const double& b3, double& x1, double& x2, double& x3) {
        const double b[3] = {b1, b2, b3};
        double x[3];
        awcm::roe_proj1(awcm::Roe1{u, H, c}, b, x);
        x1 = x[0]; x2 = x[1]; x3 = x[2];
    }
    static void Inversion(const double& u, const double& H, const double& c, const VECTOR_3D<double>& b,
                          VECTOR_3D<double>& x) {
        Inversion(u, H, c, b[0], b[1], b[2], x[0], x[1], x[2]);
    }
};

class ROE_SOLVER_2D {
public:
    // H uses gamma; the sound speed uses the literal 0.4 (ROESOLVER.h:49-71)
    static void RoeAverage(const VECTOR_4D<double>& Q_l, const VECTOR_4D<double>& Q_r, double& u, double& v, double& H,
                           double& c, const double& gamma = 1.4) {
        const awcm::Roe2 a = awcm::roe_avg2(gamma, Q_l.values, Q_r.values);
        u = a.u; v = a.v; H = a.H; c = a.c;
    }
    static void Inversion_X(const double& u, const double& v, const double& H, const double& c, const double& b1,
                            const double& b2, const double& b3, const double& b4, double& x1, double& x2, double& x3,
                            double& x4) {
        const double b[4] = {b1, b2, b3, b4};
        double x[4];
        awcm::roe_proj2<0>(awcm::Roe2{u, v, H, c}, b, x);
        x1 = x[0]; x2 = x[1]; x3 = x[2]; x4 = x[3];
    }
    static void Inversion_Y(const double& u, const double& v, const double& H, const double& c, const double& b1,
                            const double& b2, const double& b3, const double& b4, double& x1, double& x2, double& x3,
                            double& x4) {
        const double b[4] = {b1, b2, b3, b4};
        double x[4];
        awcm::roe_proj2<1>(awcm::Roe2{u, v, H, c}, b, x);
        x1 = x[0]; x2 = x[1]; x3 = x[2]; x4 = x[3];
    }
    static void Inversion_X(const double& u, const double& v, const double& H, const double& c,
                            const VECTOR_4D<double>& b, VECTOR_4D<double>& x) {
        awcm::roe_proj2<0>(awcm::Roe2{u, v, H, c}, b.values, x.values);
    }
    static void Inversion_Y(const double& u, const double& v, const double& H, const double& c,
                            const VECTOR_4D<double>& b, VECTOR_4D<double>& x) {
        awcm::roe_proj2<1>(awcm::Roe2{u, v, H, c}, b.values, x.values);
    }
};

/* ---- WENO-Z5 and WENO5-JS reconstructions and Euler face fluxes (WENOZ.h,
 *      WENO.h) -------------------------------------------------------------- */
namespace awcm_api_detail {

// smoothness indicators of the three 3-point candidate stencils
struct Beta3 {
    double b0, b1, b2;
};
inline Beta3 beta3(double ll, double l, double c, double r, double rr) {
    Beta3 s;
    s.b0 = 13.0 / 12.0 * SQR(c - 2 * r + rr) + 0.25 * SQR(3 * c - 4 * r + rr);
    s.b1 = 13.0 / 12.0 * SQR(l - 2 * c + r) + 0.25 * SQR(l - r);
    s.b2 = 13.0 / 12.0 * SQR(ll - 2 * l + c) + 0.25 * SQR(ll - 4 * l + 3 * c);
    return s;
}
// the left-biased candidate polynomials combined with normalised weights
inline double combine_left(double a0, double a1, double a2, double ll, double l, double c, double r, double rr) {
    const double as = a0 + a1 + a2;
    const double w0 = a0 / as, w1 = a1 / as, w2 = a2 / as;
    double out = w0 * (11.0 / 6.0 * c + -7.0 / 6.0 * r + 1.0 / 3.0 * rr);
    out += w1 * (1.0 / 3.0 * l + 5.0 / 6.0 * c + -1.0 / 6.0 * r);
    out += w2 * (-1.0 / 6.0 * ll + 5.0 / 6.0 * l + 1.0 / 3.0 * c);
    return out;
}
// WENO5-JS (WENO.h:71-99): a_k = d_k / (1e-6 + b_k)^2, d = (0.1, 0.6, 0.3)
inline double wenojs_left(double ll, double l, double c, double r, double rr) {
    return awcm::wenojs_left(ll, l, c, r, rr);
}
inline double wenojs_right(double ll, double l, double c, double r, double rr) { return wenojs_left(rr, r, c, l, ll); }

// Roe-characteristic Lax-Friedrichs split flux at the face s[2]|s[3] with
// WENO5-JS reconstructions (WENO.h:129-174, 216-311): the device's
// awcm::face_flux1/face_flux2 with the WENO5-JS reconstruction.
inline void face_flux1_js(double g, const double (*s)[3], double alpha, double* flux) {
    awcm::face_flux1<awcm::kRecWeno5>(g, s, alpha, flux);
}
template <int DIR>
inline void face_flux2_js(double g, const double (*s)[4], double alpha, double* flux) {
    awcm::face_flux2<DIR, awcm::kRecWeno5>(g, s, alpha, flux);
}

template <class V, int N>
inline void stack6(const V& a, const V& b, const V& c, const V& d, const V& e, const V& f, double (*s)[N]) {
    const V* v[6] = {&a, &b, &c, &d, &e, &f};
    for (int p = 0; p < 6; p++)
        for (int k = 0; k < N; k++) s[p][k] = (*v[p])[k];
}

}  // namespace awcm_api_detail

class WENOZ5 {
public:
    double eps = 1e-40;
    void ReconstructLeft(const double& ll, const double& l, const double& c, const double& r, const double& rr,
                         double& left) const {
        left = awcm::wenoz_left(ll, l, c, r, rr);
    }
    void ReconstructRight(const double& ll, const double& l, const double& c, const double& r, const double& rr,
                          double& right) const {
        right = awcm::wenoz_right(ll, l, c, r, rr);
    }
    // WENOZ.h:29-73 computes `right` with Z weights d = (0.3, 0.6, 0.1) on the
    // right-biased polynomials and `left` with JS-form weights d_k / (eps +
    // b_k)^2 at eps = 1e-40; reproduced as written.
    void Reconstruct(const double& ll, const double& l, const double& c, const double& r, const double& rr, double& left,
                     double& right) const {
        const awcm_api_detail::Beta3 s = awcm_api_detail::beta3(ll, l, c, r, rr);
        const double tau = ABS(s.b0 - s.b2);
        const double a0 = 0.3 * (1 + tau / (s.b0 + eps)), a1 = 0.6 * (1 + tau / (s.b1 + eps)),
                     a2 = 0.1 * (1 + tau / (s.b2 + eps));
        const double as = a0 + a1 + a2;
        const double w0 = a0 / as, w1 = a1 / as, w2 = a2 / as;
        right = w0 * (1.0 / 3.0 * c + 5.0 / 6.0 * r + -1.0 / 6.0 * rr);
        right += w1 * (-1.0 / 6.0 * l + 5.0 / 6.0 * c + 1.0 / 3.0 * r);
        right += w2 * (1.0 / 3.0 * ll + -7.0 / 6.0 * l + 11.0 / 6.0 * c);
        left = awcm_api_detail::combine_left(0.1 / SQR(eps + s.b0), 0.6 / SQR(eps + s.b1), 0.3 / SQR(eps + s.b2), ll, l, c,
                                             r, rr);
    }
};

class WENO5 {
public:
    double eps = 1e-6;
    void ReconstructLeft(const double& ll, const double& l, const double& c, const double& r, const double& rr,
                         double& left) const {
        left = awcm_api_detail::wenojs_left(ll, l, c, r, rr);
    }
    void ReconstructRight(const double& ll, const double& l, const double& c, const double& r, const double& rr,
                          double& right) const {
        right = awcm_api_detail::wenojs_right(ll, l, c, r, rr);
    }
    // WENO.h:28-69: `right` with d = (0.3, 0.6, 0.1) on the right-biased
    // polynomials, `left` as ReconstructLeft
    void Reconstruct(const double& ll, const double& l, const double& c, const double& r, const double& rr, double& left,
                     double& right) const {
        const awcm_api_detail::Beta3 s = awcm_api_detail::beta3(ll, l, c, r, rr);
        const double a0 = 0.3 / SQR(eps + s.b0), a1 = 0.6 / SQR(eps + s.b1), a2 = 0.1 / SQR(eps + s.b2);
        const double as = a0 + a1 + a2;
        const double w0 = a0 / as, w1 = a1 / as, w2 = a2 / as;
        right = w0 * (1.0 / 3.0 * c + 5.0 / 6.0 * r + -1.0 / 6.0 * rr);
        right += w1 * (-1.0 / 6.0 * l + 5.0 / 6.0 * c + 1.0 / 3.0 * r);
        right += w2 * (1.0 / 3.0 * ll + -7.0 / 6.0 * l + 11.0 / 6.0 * c);
        left = awcm_api_detail::combine_left(0.1 / SQR(eps + s.b0), 0.6 / SQR(eps + s.b1), 0.3 / SQR(eps + s.b2), ll, l, c,
                                             r, rr);
    }
};

class WENOZ_EULER_SOLVER_1D {
public:
    WENOZ5 wenoz;
    const double gamma;
    WENOZ_EULER_SOLVER_1D(const double& gamma_ = 1.4) : gamma(gamma_) {}
    void ComputeFlux(const VECTOR_3D<double>& lll, const VECTOR_3D<double>& ll, const VECTOR_3D<double>& l,
                     const VECTOR_3D<double>& r, const VECTOR_3D<double>& rr, const VECTOR_3D<double>& rrr,
                     const double& alpha, VECTOR_3D<double>& flux) const {
        double s[6][3];
        awcm_api_detail::stack6<VECTOR_3D<double>, 3>(lll, ll, l, r, rr, rrr, s);
        awcm::face_flux1(gamma, s, alpha, flux.values);
    }
};

class WENO_EULER_SOLVER_1D {
public:
    WENO5 weno;
    const double gamma;
    WENO_EULER_SOLVER_1D(const double& gamma_ = 1.4) : gamma(gamma_) {}
    void ComputeFlux(const VECTOR_3D<double>& lll, const VECTOR_3D<double>& ll, const VECTOR_3D<double>& l,
                     const VECTOR_3D<double>& r, const VECTOR_3D<double>& rr, const VECTOR_3D<double>& rrr,
                     const double& alpha, VECTOR_3D<double>& flux) const {
        double s[6][3];
        awcm_api_detail::stack6<VECTOR_3D<double>, 3>(lll, ll, l, r, rr, rrr, s);
        awcm_api_detail::face_flux1_js(gamma, s, alpha, flux.values);
    }
    // dF = F(c+1/2) - F(c-1/2), not divided by dx (WENO.h:176-184)
    void ComputedF(const VECTOR_3D<double>& lll, const VECTOR_3D<double>& ll, const VECTOR_3D<double>& l,
                   const VECTOR_3D<double>& c, const VECTOR_3D<double>& r, const VECTOR_3D<double>& rr,
                   const VECTOR_3D<double>& rrr, const double& alpha, VECTOR_3D<double>& dF) const {
        VECTOR_3D<double> F_l, F_r;
        ComputeFlux(lll, ll, l, c, r, rr, alpha, F_l);
        ComputeFlux(ll, l, c, r, rr, rrr, alpha, F_r);
        dF = F_r - F_l;
    }
};

class WENOZ_EULER_SOLVER_2D {
public:
    WENOZ5 wenoz;
    const double gamma;
    WENOZ_EULER_SOLVER_2D(const double& gamma_ = 1.4) : gamma(gamma_) {}
    void ComputeFluxX(const VECTOR_4D<double>& lll, const VECTOR_4D<double>& ll, const VECTOR_4D<double>& l,
                      const VECTOR_4D<double>& r, const VECTOR_4D<double>& rr, const VECTOR_4D<double>& rrr,
                      const double& alpha, VECTOR_4D<double>& flux) const {
        double s[6][4];
        awcm_api_detail::stack6<VECTOR_4D<double>, 4>(lll, ll, l, r, rr, rrr, s);
        awcm::face_flux2<0>(gamma, s, alpha, flux.values);
    }
    void ComputeFluxY(const VECTOR_4D<double>& bbb, const VECTOR_4D<double>& bb, const VECTOR_4D<double>& b,
                      const VECTOR_4D<double>& t, const VECTOR_4D<double>& tt, const VECTOR_4D<double>& ttt,
                      const double& alpha, VECTOR_4D<double>& flux) const {
        double s[6][4];
        awcm_api_detail::stack6<VECTOR_4D<double>, 4>(bbb, bb, b, t, tt, ttt, s);
        awcm::face_flux2<1>(gamma, s, alpha, flux.values);
    }
};

class WENO_EULER_SOLVER_2D {
public:
    WENO5 weno;
    const double gamma;
    WENO_EULER_SOLVER_2D(const double& gamma_ = 1.4) : gamma(gamma_) {}
    void ComputeFluxX(const VECTOR_4D<double>& lll, const VECTOR_4D<double>& ll, const VECTOR_4D<double>& l,
                      const VECTOR_4D<double>& r, const VECTOR_4D<double>& rr, const VECTOR_4D<double>& rrr,
                      const double& alpha, VECTOR_4D<double>& flux) const {
        double s[6][4];
        awcm_api_detail::stack6<VECTOR_4D<double>, 4>(lll, ll, l, r, rr, rrr, s);
        awcm_api_detail::face_flux2_js<0>(gamma, s, alpha, flux.values);
    }
    void ComputeFluxY(const VECTOR_4D<double>& bbb, const VECTOR_4D<double>& bb, const VECTOR_4D<double>& b,
                      const VECTOR_4D<double>& t, const VECTOR_4D<double>& tt, const VECTOR_4D<double>& ttt,
                      const double& alpha, VECTOR_4D<double>& flux) const {
        double s[6][4];
        awcm_api_detail::stack6<VECTOR_4D<double>, 4>(bbb, bb, b, t, tt, ttt, s);
        awcm_api_detail::face_flux2_js<1>(gamma, s, alpha, flux.values);
    }
};

/* ---- OMP_MANAGER.h, STOP_WATCH.h ----------------------------------------- */
class OMP_MANAGER {
public:
    int num_thread = 0;
    double* arr_double = nullptr;
    int* arr_int = nullptr;
    OMP_MANAGER() {}
    OMP_MANAGER(const int& n) { Initialize(n); }
    ~OMP_MANAGER() {
        SAFE_DELETE_ARRAY(arr_double);
        SAFE_DELETE_ARRAY(arr_int);
    }
    OMP_MANAGER(const OMP_MANAGER&) = delete;
    OMP_MANAGER& operator=(const OMP_MANAGER&) = delete;
    // slots start at zero (the reference leaves them unset), so a team
    // smaller than num_thread reduces over zeros rather than garbage
    void Initialize(const int& n) {
        num_thread = n;
        SAFE_DELETE_ARRAY(arr_double);
        SAFE_DELETE_ARRAY(arr_int);
        arr_double = new double[num_thread]();
        arr_int = new int[num_thread]();
#ifdef _OPENMP
        omp_set_num_threads(num_thread);
#endif
    }
    static int thread_id() {
#ifdef _OPENMP
        return omp_get_thread_num();
#else
        return 0;
#endif
    }
    void Input(const double& v) { arr_double[thread_id()] = v; }
    void Input(const int& v) { arr_int[thread_id()] = v; }
    double Max_double() {
        double v = arr_double[0];
        for (int k = 1; k < num_thread; k++) v = MAX(v, arr_double[k]);
        return v;
    }
    int Max_int() {
        int v = arr_int[0];
        for (int k = 1; k < num_thread; k++) v = MAX(v, arr_int[k]);
        return v;
    }
    double Min_double() {
        double v = arr_double[0];
        for (int k = 1; k < num_thread; k++) v = MIN(v, arr_double[k]);
        return v;
    }
    int Min_int() {
        int v = arr_int[0];
        for (int k = 1; k < num_thread; k++) v = MIN(v, arr_int[k]);
        return v;
    }
    double Sum_double() {
        double v = arr_double[0];
        for (int k = 1; k < num_thread; k++) v += arr_double[k];
        return v;
    }
    int Sum_int() {
        int v = arr_int[0];
        for (int k = 1; k < num_thread; k++) v += arr_int[k];
        return v;
    }
};

class STOP_WATCH {
public:
    std::clock_t start_time = 0, stop_time = 0;
    int num_toc = 0;
    void Tic() {
        start_time = std::clock();
        num_toc = 0;
    }
    double Toc() {
        stop_time = std::clock();
        num_toc++;
        return double(stop_time - start_time) / CLOCKS_PER_SEC;
    }
    void TocAndReadDuration(const char* name_work) { std::printf("%2.3f seconds in %s", Toc(), name_work); }
};

/* ---- AWCM managers (AWCM_MANAGER_1D.h, AWCM_MANAGER_2D.h): the static mask
 *      algebra on the GPU, and the output writers --------------------------- */
namespace awcm_api_detail {
// The static members run on a device solver of the MRA grid's geometry, one
// per geometry, created on first use and held until the process ends; calls
// on it are serialised.
template <class F>
void on_grid(int dim, int L, int nxc, int nyc, F&& f) {
    static std::mutex mu;
    static auto* cache = new std::map<std::vector<int>, awcm_solver*>();
    std::lock_guard<std::mutex> lock(mu);
    const std::vector<int> key{dim, L, nxc, nyc};
    auto it = cache->find(key);
    awcm_solver* h = it == cache->end() ? nullptr : it->second;
    if (!h) {
        const int st = dim == 1 ? awcm_create_1d(L, nxc, 1.4, 1.0, &h)
                                : awcm_create_2d(L, nxc, nyc, 1.4, 1.0, 0.0, AWCM_BC_TRANSMISSIVE, &h);
        if (st != AWCM_OK)
            throw std::runtime_error("AWCM_MANAGER: device solver creation failed with AWCM status " + std::to_string(st));
        (*cache)[key] = h;
    }
    f(h);
}
inline void manager_check(int st, const char* what) {
    if (st != AWCM_OK) throw std::runtime_error(std::string(what) + " failed with AWCM status " + std::to_string(st));
}
// A mask or field argument spans the finest grid with any ghost width; it
// crosses the C ABI with ghost width 3 (ghosts it lacks zero) and comes back
// over its own extent (points the GPU set beyond it, where the reference
// would write out of bounds, are dropped).
template <class F>
void span_check_1d(const F& f, int Nx, const char* who) {
    if (f.i_start != 0 || f.i_res != Nx) throw std::invalid_argument(std::string(who) + ": fields must span 0..Nx-1");
}
template <class F>
void span_check_2d(const F& f, int Nx, int Ny, const char* who) {
    if (f.i_start != 0 || f.j_start != 0 || f.i_res != Nx || f.j_res != Ny)
        throw std::invalid_argument(std::string(who) + ": fields must span (0..Nx-1) x (0..Ny-1)");
}
template <class F>
int ghosts_in(const F& f) { return f.ghost_width < 3 ? f.ghost_width : 3; }
}  // namespace awcm_api_detail

class AWCM_MANAGER_1D {
public:
    // AddAdjacentZone (AWCM_MANAGER_1D.h:14-43)
    static void AddAdjacentZone(FIELD_UNIFORM_1D<bool>& mask, const MRA_GRID_INT_1D& mra) {
        mask_op(0, mask, nullptr, mra, "AWCM_MANAGER_1D::AddAdjacentZone");
    }
    // CheckReconstruction (AWCM_MANAGER_1D.h:45-71)
    static void CheckReconstruction(FIELD_UNIFORM_1D<bool>& mask, const MRA_GRID_INT_1D& mra) {
        mask_op(1, mask, nullptr, mra, "AWCM_MANAGER_1D::CheckReconstruction");
    }
    // AddPointForDifferentiation (AWCM_MANAGER_1D.h:73-107)
    static void AddPointForDifferentiation(FIELD_UNIFORM_1D<bool>& mask, const MRA_GRID_INT_1D& mra,
                                           const FIELD_UNIFORM_1D<int>& level) {
        mask_op(2, mask, const_cast<FIELD_UNIFORM_1D<int>*>(&level), mra, "AWCM_MANAGER_1D::AddPointForDifferentiation");
    }
    // ComputeLevel (AWCM_MANAGER_1D.h:109-133)
    static void ComputeLevel(FIELD_UNIFORM_1D<int>& level, const FIELD_UNIFORM_1D<bool>& mask,
                             const MRA_GRID_INT_1D& mra) {
        mask_op(3, const_cast<FIELD_UNIFORM_1D<bool>&>(mask), &level, mra, "AWCM_MANAGER_1D::ComputeLevel");
    }
    // RemoveResidue (AWCM_MANAGER_1D.h:135-144, 146-157)
    static void RemoveResidue(FIELD_UNIFORM_1D<double>& q, const FIELD_UNIFORM_1D<bool>& mask,
                              const MRA_GRID_INT_1D& mra) {
        residue(q, 1, [](FIELD_UNIFORM_1D<double>& f, int i, int) -> double& { return f(i); }, mask, mra);
    }
    static void RemoveResidue(FIELD_UNIFORM_1D<VECTOR_3D<double>>& q, const FIELD_UNIFORM_1D<bool>& mask,
                              const MRA_GRID_INT_1D& mra) {
        residue(q, 3, [](FIELD_UNIFORM_1D<VECTOR_3D<double>>& f, int i, int c) -> double& { return f(i)[c]; }, mask,
                mra);
    }

    template <class T>
    static void Print(const FIELD_UNIFORM_1D<T>& q, const FIELD_UNIFORM_1D<bool>& mask, const char* filename) {
        std::ofstream f(filename);
        const GRID_UNIFORM_1D& grid(q.base_grid);
        ITERATION_GRID_1D(q, 0) {
            if (mask(i) == true) f << grid(i) << " " << q(i) << std::endl;
        }
    }

private:
    static std::vector<uint8_t> mask_in(const FIELD_UNIFORM_1D<bool>& m, int Nx) {
        std::vector<uint8_t> b(size_t(Nx) + 6, 0);
        const int g = awcm_api_detail::ghosts_in(m);
        for (int i = -g; i <= Nx - 1 + g; i++) b[size_t(i + 3)] = m(i) ? 1 : 0;
        return b;
    }
    static void mask_op(int op, FIELD_UNIFORM_1D<bool>& mask, FIELD_UNIFORM_1D<int>* level,
                        const MRA_GRID_INT_1D& mra, const char* who) {
        const int Nx = mra.Nx;
        awcm_api_detail::span_check_1d(mask, Nx, who);
        if (level) awcm_api_detail::span_check_1d(*level, Nx, who);
        std::vector<uint8_t> b = mask_in(mask, Nx);
        std::vector<int32_t> lv(size_t(Nx), 0);
        if (level)
            for (int i = 0; i < Nx; i++) lv[size_t(i)] = (*level)(i);
        awcm_api_detail::on_grid(1, mra.num_levels, mra.Nx_coarse, 0, [&](awcm_solver* h) {
            int st = AWCM_OK;
            if (op == 0) st = awcm_add_adjacent_zone(h, b.data());
            else if (op == 1) st = awcm_check_reconstruction(h, b.data());
            else if (op == 2) st = awcm_add_point_for_differentiation(h, b.data(), lv.data());
            else st = awcm_compute_level(h, lv.data(), b.data());
            awcm_api_detail::manager_check(st, who);
        });
        if (op == 3) {
            for (int i = 0; i < Nx; i++) (*level)(i) = lv[size_t(i)];
            return;
        }
        const int g = awcm_api_detail::ghosts_in(mask);
        for (int i = -g; i <= Nx - 1 + g; i++) mask(i) = b[size_t(i + 3)] != 0;
    }
    template <class Field, class At>
    static void residue(Field& q, int nc, At at, const FIELD_UNIFORM_1D<bool>& mask, const MRA_GRID_INT_1D& mra) {
        const char* who = "AWCM_MANAGER_1D::RemoveResidue";
        const int Nx = mra.Nx, g = awcm_api_detail::ghosts_in(q);
        awcm_api_detail::span_check_1d(q, Nx, who);
        awcm_api_detail::span_check_1d(mask, Nx, who);
        std::vector<double> v((size_t(Nx) + 6) * nc, 0.0);
        for (int i = -g; i <= Nx - 1 + g; i++)
            for (int c = 0; c < nc; c++) v[size_t(i + 3) * nc + c] = at(q, i, c);
        std::vector<uint8_t> b = mask_in(mask, Nx);
        awcm_api_detail::on_grid(1, mra.num_levels, mra.Nx_coarse, 0, [&](awcm_solver* h) {
            awcm_api_detail::manager_check(awcm_remove_residue(h, v.data(), nc, b.data()), who);
        });
        for (int i = -g; i <= Nx - 1 + g; i++)
            for (int c = 0; c < nc; c++) at(q, i, c) = v[size_t(i + 3) * nc + c];
    }
};

class AWCM_MANAGER_2D {
public:
    // AddAdjacentZone (AWCM_MANAGER_2D.h:14-61)
    static void AddAdjacentZone(FIELD_UNIFORM_2D<bool>& mask, const MRA_GRID_INT_2D& mra_grid) {
        mask_op(0, mask, nullptr, mra_grid, "AWCM_MANAGER_2D::AddAdjacentZone");
    }
    // CheckReconstruction (AWCM_MANAGER_2D.h:63-158)
    static void CheckReconstruction(FIELD_UNIFORM_2D<bool>& mask, const MRA_GRID_INT_2D& mra_grid) {
        mask_op(1, mask, nullptr, mra_grid, "AWCM_MANAGER_2D::CheckReconstruction");
    }
    // AddPointForDifferentiation (AWCM_MANAGER_2D.h:160-213)
    static void AddPointForDifferentiation(FIELD_UNIFORM_2D<bool>& mask, const MRA_GRID_INT_2D& mra_grid,
                                           const FIELD_UNIFORM_2D<int>& level) {
        mask_op(2, mask, const_cast<FIELD_UNIFORM_2D<int>*>(&level), mra_grid,
                "AWCM_MANAGER_2D::AddPointForDifferentiation");
    }
    // ComputeLevel (AWCM_MANAGER_2D.h:215-241)
    static void ComputeLevel(FIELD_UNIFORM_2D<int>& level, const FIELD_UNIFORM_2D<bool>& mask,
                             const MRA_GRID_INT_2D& mra_grid) {
        mask_op(3, const_cast<FIELD_UNIFORM_2D<bool>&>(mask), &level, mra_grid, "AWCM_MANAGER_2D::ComputeLevel");
    }
    // RemoveResidue (AWCM_MANAGER_2D.h:257-266)
    static void RemoveResidue(FIELD_UNIFORM_2D<double>& q, const FIELD_UNIFORM_2D<bool>& mask,
                              const MRA_GRID_INT_2D& mra_grid) {
        residue(q, 1, [](FIELD_UNIFORM_2D<double>& f, int i, int j, int) -> double& { return f(i, j); }, mask,
                mra_grid, "AWCM_MANAGER_2D::RemoveResidue");
    }
    // RemoveResidueEuler (AWCM_MANAGER_2D.h:267-280)
    static void RemoveResidueEuler(FIELD_UNIFORM_2D<VECTOR_4D<double>>& q, const FIELD_UNIFORM_2D<bool>& mask,
                                   const MRA_GRID_INT_2D& mra_grid) {
        residue(q, 4, [](FIELD_UNIFORM_2D<VECTOR_4D<double>>& f, int i, int j, int c) -> double& { return f(i, j)[c]; },
                mask, mra_grid, "AWCM_MANAGER_2D::RemoveResidueEuler");
    }

    template <class T>
    static void Print(const FIELD_UNIFORM_2D<T>& q, const FIELD_UNIFORM_2D<bool>& mask, const char* filename) {
        std::ofstream f(filename);
        f.precision(15);
        const GRID_UNIFORM_2D& grid(q.grid);
        ITERATION_GRID_2D(q, 0) {
            if (mask(i, j) == false) continue;
            f << grid(i, j).i << " " << grid(i, j).j << " " << q(i, j) << std::endl;
        }
    }

private:
    static std::vector<uint8_t> mask_in(const FIELD_UNIFORM_2D<bool>& m, int Nx, int Ny) {
        const size_t P = size_t(Nx) + 6;
        std::vector<uint8_t> b(P * (size_t(Ny) + 6), 0);
        const int g = awcm_api_detail::ghosts_in(m);
        for (int j = -g; j <= Ny - 1 + g; j++)
            for (int i = -g; i <= Nx - 1 + g; i++) b[size_t(i + 3) + P * size_t(j + 3)] = m(i, j) ? 1 : 0;
        return b;
    }
    static void mask_op(int op, FIELD_UNIFORM_2D<bool>& mask, FIELD_UNIFORM_2D<int>* level,
                        const MRA_GRID_INT_2D& mra_grid, const char* who) {
        const int Nx = mra_grid.Nx, Ny = mra_grid.Ny;
        awcm_api_detail::span_check_2d(mask, Nx, Ny, who);
        if (level) awcm_api_detail::span_check_2d(*level, Nx, Ny, who);
        std::vector<uint8_t> b = mask_in(mask, Nx, Ny);
        std::vector<int32_t> lv(size_t(Nx) * Ny, 0);
        if (level)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) lv[size_t(i) + size_t(Nx) * j] = (*level)(i, j);
        awcm_api_detail::on_grid(2, mra_grid.num_levels, mra_grid.Nx_coarse, mra_grid.Ny_coarse, [&](awcm_solver* h) {
            int st = AWCM_OK;
            if (op == 0) st = awcm_add_adjacent_zone(h, b.data());
            else if (op == 1) st = awcm_check_reconstruction(h, b.data());
            else if (op == 2) st = awcm_add_point_for_differentiation(h, b.data(), lv.data());
            else st = awcm_compute_level(h, lv.data(), b.data());
            awcm_api_detail::manager_check(st, who);
        });
        if (op == 3) {
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) (*level)(i, j) = lv[size_t(i) + size_t(Nx) * j];
            return;
        }
        const size_t P = size_t(Nx) + 6;
        const int g = awcm_api_detail::ghosts_in(mask);
        for (int j = -g; j <= Ny - 1 + g; j++)
            for (int i = -g; i <= Nx - 1 + g; i++) mask(i, j) = b[size_t(i + 3) + P * size_t(j + 3)] != 0;
    }
    template <class Field, class At>
    static void residue(Field& q, int nc, At at, const FIELD_UNIFORM_2D<bool>& mask, const MRA_GRID_INT_2D& mra_grid,
                        const char* who) {
        const int Nx = mra_grid.Nx, Ny = mra_grid.Ny, g = awcm_api_detail::ghosts_in(q);
        awcm_api_detail::span_check_2d(q, Nx, Ny, who);
        awcm_api_detail::span_check_2d(mask, Nx, Ny, who);
        const size_t P = size_t(Nx) + 6;
        std::vector<double> v(P * (size_t(Ny) + 6) * nc, 0.0);
        for (int j = -g; j <= Ny - 1 + g; j++)
            for (int i = -g; i <= Nx - 1 + g; i++)
                for (int c = 0; c < nc; c++) v[(size_t(i + 3) + P * size_t(j + 3)) * nc + c] = at(q, i, j, c);
        std::vector<uint8_t> b = mask_in(mask, Nx, Ny);
        awcm_api_detail::on_grid(2, mra_grid.num_levels, mra_grid.Nx_coarse, mra_grid.Ny_coarse, [&](awcm_solver* h) {
            awcm_api_detail::manager_check(awcm_remove_residue(h, v.data(), nc, b.data()), who);
        });
        for (int j = -g; j <= Ny - 1 + g; j++)
            for (int i = -g; i <= Nx - 1 + g; i++)
                for (int c = 0; c < nc; c++) at(q, i, j, c) = v[(size_t(i + 3) + P * size_t(j + 3)) * nc + c];
    }
};

/* ---- AWCM_1D / AWCM_2D ---------------------------------------------------- */
namespace awcm_api_detail {
inline void check(int st, const char* what) {
    if (st != AWCM_OK) throw std::runtime_error(std::string(what) + " failed with AWCM status " + std::to_string(st));
}
// one device solver per AWCM object, created by the first PrepareEuler
struct Handle {
    awcm_solver* h = nullptr;
    Handle() = default;
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    ~Handle() {
        if (h) awcm_destroy(h);
    }
};
// After a prepare: mask_temp, mask0..mask3 arrive ghosted in the fields' own
// layout, level0..level3 unghosted.
template <class MaskField, class LevelField>
void fetch_masks(awcm_solver* h, size_t n_masked, size_t n_levels, MaskField* const (&ms)[5],
                 LevelField* const (&ls)[4]) {
    std::unique_ptr<uint8_t[]> m(new uint8_t[5 * n_masked]);
    std::unique_ptr<int32_t[]> lv(new int32_t[4 * n_levels]);
    check(awcm_get_masks(h, m.get(), lv.get()), "awcm_get_masks");
    for (int k = 0; k < 5; k++)
        for (size_t n = 0; n < n_masked; n++) ms[k]->arr.val[n] = m[k * n_masked + n] != 0;
    for (int k = 0; k < 4; k++)
        for (size_t n = 0; n < n_levels; n++) ls[k]->arr.val[n] = lv[k * n_levels + n];
}
template <class MaskField>
void push_mask0(awcm_solver* h, const MaskField& mask0, size_t n_masked) {
    std::unique_ptr<uint8_t[]> m(new uint8_t[n_masked]);
    for (size_t n = 0; n < n_masked; n++) m[n] = mask0.arr.val[n] ? 1 : 0;
    check(awcm_set_mask0(h, m.get()), "awcm_set_mask0");
}
}  // namespace awcm_api_detail

namespace awcm_api_detail {
// the 1-D mask argument in the C ABI's layout, Nx+6 bytes (ghosts it lacks
// zero); only its finest-grid points are read, as in the reference
inline std::unique_ptr<uint8_t[]> mask_bytes(const FIELD_UNIFORM_1D<bool>& m, int Nx) {
    if (m.i_start != 0 || m.i_res != Nx) throw std::invalid_argument("WAVELET_1D: masks must span 0..Nx-1");
    const int g = m.ghost_width < 3 ? m.ghost_width : 3;
    std::unique_ptr<uint8_t[]> b(new uint8_t[size_t(Nx) + 6]());
    for (int i = -g; i <= Nx - 1 + g; i++) b[size_t(i + 3)] = m(i) ? 1 : 0;
    return b;
}
}  // namespace awcm_api_detail

// WAVELET_1D (Wavelet/WAVELET_1D.h:11-42): N and weight(i, j) as in WAVELET_2D;
// the interpolations (WAVELET_1D.h:198-275, 393-473) run on the GPU, on the
// owning AWCM_1D's device solver (a free-standing one has its own, created on
// first use from the field's grid).
class WAVELET_1D {
public:
    const int N;
    ARRAY_2D<double> weight;

    WAVELET_1D(const int& N_ = 2) : N(N_) {
        if (N != 2) throw std::invalid_argument("WAVELET_1D: only N = 2 (the 4-point wavelet) is supported");
        weight.Initialize(0, -1, 2 * N, 2 * N + 1);
        for (int j = -1; j < 2 * N; j++)
            for (int i = 0; i < 2 * N; i++) {
                double l = 1.0;
                for (int k = 0; k < 2 * N; k++)
                    if (k != i) l *= double(j + 0.5 - k) / double(i - k);
                weight(i, j) = l;
            }
    }
    WAVELET_1D(const WAVELET_1D&) = delete;
    WAVELET_1D& operator=(const WAVELET_1D&) = delete;

    // WAVELET_1D::Interpolate(q, mra, mask_from, mask_to) (WAVELET_1D.h:198-233)
    void Interpolate(FIELD_UNIFORM_1D<double>& q, const MRA_GRID_INT_1D& mra, const FIELD_UNIFORM_1D<bool>& mask_from,
                     const FIELD_UNIFORM_1D<bool>& mask_to) const {
        run(q, mra, awcm_api_detail::mask_bytes(mask_from, mra.Nx).get(), awcm_api_detail::mask_bytes(mask_to, mra.Nx).get());
    }
    // WAVELET_1D::Interpolate(q, mra, mask_from) (WAVELET_1D.h:235-275): every
    // point not in mask_from
    void Interpolate(FIELD_UNIFORM_1D<double>& q, const MRA_GRID_INT_1D& mra,
                     const FIELD_UNIFORM_1D<bool>& mask_from) const {
        std::unique_ptr<uint8_t[]> all(new uint8_t[size_t(mra.Nx) + 6]);
        for (int i = 0; i < mra.Nx + 6; i++) all[size_t(i)] = 1;
        run(q, mra, awcm_api_detail::mask_bytes(mask_from, mra.Nx).get(), all.get());
    }

    // WAVELET_1D::InterpolateEuler(q, mra, mask_from, mask_to)
    // (WAVELET_1D.h:393-431): momentum and energy
    void InterpolateEuler(FIELD_UNIFORM_1D<VECTOR_3D<double>>& q, const MRA_GRID_INT_1D& mra,
                          const FIELD_UNIFORM_1D<bool>& mask_from, const FIELD_UNIFORM_1D<bool>& mask_to) const {
        using awcm_api_detail::check;
        awcm_solver* h = open(q.ghost_width, q.i_start, q.i_res, q.base_grid.dx, mra);
        check(awcm_upload_q(h, q.arr.val[0].values), "awcm_upload_q");
        check(awcm_interpolate_euler(h, awcm_api_detail::mask_bytes(mask_from, mra.Nx).get(),
                                     awcm_api_detail::mask_bytes(mask_to, mra.Nx).get()),
              "awcm_interpolate_euler");
        check(awcm_download_q(h, q.arr.val[0].values), "awcm_download_q");
    }
    // WAVELET_1D::InterpolateEuler(q, mra, mask_from) (WAVELET_1D.h:433-473):
    // every point not in mask_from, all three components
    void InterpolateEuler(FIELD_UNIFORM_1D<VECTOR_3D<double>>& q, const MRA_GRID_INT_1D& mra,
                          const FIELD_UNIFORM_1D<bool>& mask_from) const {
        using awcm_api_detail::check;
        awcm_solver* h = open(q.ghost_width, q.i_start, q.i_res, q.base_grid.dx, mra);
        check(awcm_upload_q(h, q.arr.val[0].values), "awcm_upload_q");
        check(awcm_interpolate_all_grid(h, awcm_api_detail::mask_bytes(mask_from, mra.Nx).get()),
              "awcm_interpolate_all_grid");
        check(awcm_download_q(h, q.arr.val[0].values), "awcm_download_q");
    }

private:
    friend class AWCM_1D;
    mutable awcm_api_detail::Handle own;
    awcm_api_detail::Handle* dev = &own;
    awcm_solver* open(int ghost, int i_start, int i_res, double dx, const MRA_GRID_INT_1D& mra) const {
        using awcm_api_detail::check;
        if (ghost != 3 || i_start != 0 || i_res != mra.Nx)
            throw std::invalid_argument("WAVELET_1D: the field must span 0..Nx-1 with ghost width 3");
        if (!dev->h) check(awcm_create_1d(mra.num_levels, mra.Nx_coarse, 1.4, dx, &dev->h), "awcm_create_1d");
        return dev->h;
    }
    void run(FIELD_UNIFORM_1D<double>& q, const MRA_GRID_INT_1D& mra, const uint8_t* from, const uint8_t* to) const {
        using awcm_api_detail::check;
        awcm_solver* h = open(q.ghost_width, q.i_start, q.i_res, q.base_grid.dx, mra);
        check(awcm_upload_scalar(h, q.arr.val), "awcm_upload_scalar");
        check(awcm_interpolate(h, from, to), "awcm_interpolate");
        check(awcm_download_scalar(h, q.arr.val), "awcm_download_scalar");
    }
};

class AWCM_1D {
public:
    WAVELET_1D wavelet;
    MRA_GRID_INT_1D& mra;
    FIELD_UNIFORM_1D<bool> mask_temp, mask0, mask1, mask2, mask3;
    FIELD_UNIFORM_1D<int> level0, level1, level2, level3;
    const double d1, d2;
    double d00, d01, d02;
    double d10, d11, d12, d13;
    double d20, d21, d22, d23, d24;
    double d30, d31, d32, d33, d34, d35;
    awcm_api_detail::Handle dev;

    AWCM_1D(MRA_GRID_INT_1D& mra_) : mra(mra_), d1(2.0 / 3.0), d2(-1.0 / 12.0) {
        wavelet.dev = &dev;
        FIELD_UNIFORM_1D<bool>* ms[5] = {&mask_temp, &mask0, &mask1, &mask2, &mask3};
        FIELD_UNIFORM_1D<int>* ls[4] = {&level0, &level1, &level2, &level3};
        for (auto* m : ms) m->Initialize(0, mra.Nx, 0, mra.Nx, 3);
        for (auto* l : ls) l->Initialize(0, mra.Nx, 0, mra.Nx);
        mask0.AssignAllValues(true);
        d00 = -11.0 / 6.0; d01 = -1.0 / 3.0; d02 = 1.0 / 12.0;
        d10 = 3.0; d11 = -1.0 / 2.0; d12 = -2.0 / 3.0; d13 = 1.0 / 12.0;
        d20 = -3.0 / 2.0; d21 = 1.0; d22 = 0; d23 = -2.0 / 3.0; d24 = 1.0 / 12.0;
        d30 = 1.0 / 3.0; d31 = -1.0 / 6.0; d32 = 2.0 / 3.0; d33 = 0; d34 = -2.0 / 3.0; d35 = 1.0 / 12.0;
    }

    // AWCM_1D::Prepare (AWCM_1D.h:45-71) on the GPU: the scalar path.  mask0
    // and the field go down; the field, mask0..mask3 and level0..level3 come
    // back (mask_temp is not the scalar path's and stays as it was).
    void Prepare(FIELD_UNIFORM_1D<double>& q, const double& tol = 1e-6) {
        using awcm_api_detail::check;
        if (q.ghost_width != 3 || q.i_start != 0 || q.i_res != mra.Nx)
            throw std::invalid_argument("AWCM_1D::Prepare: the field must span 0..Nx-1 with ghost width 3");
        if (!dev.h) check(awcm_create_1d(mra.num_levels, mra.Nx_coarse, 1.4, q.base_grid.dx, &dev.h), "awcm_create_1d");
        const size_t NP = size_t(mra.Nx) + 6;
        awcm_api_detail::push_mask0(dev.h, mask0, NP);
        check(awcm_upload_scalar(dev.h, q.arr.val), "awcm_upload_scalar");
        check(awcm_prepare(dev.h, tol), "awcm_prepare");
        check(awcm_download_scalar(dev.h, q.arr.val), "awcm_download_scalar");
        // (FIELD_UNIFORM_1D's copy constructor drops the ghosts, so the values
        // are kept as bytes)
        std::vector<uint8_t> keep(NP);
        for (size_t n = 0; n < NP; n++) keep[n] = mask_temp.arr.val[n];
        FIELD_UNIFORM_1D<bool>* const ms[5] = {&mask_temp, &mask0, &mask1, &mask2, &mask3};
        FIELD_UNIFORM_1D<int>* const ls[4] = {&level0, &level1, &level2, &level3};
        awcm_api_detail::fetch_masks(dev.h, NP, size_t(mra.Nx), ms, ls);
        for (size_t n = 0; n < NP; n++) mask_temp.arr.val[n] = keep[n] != 0;
    }
    // AWCM_1D.h:73-80: the 3-wide band at both ends is always active
    void BadingBoundary(FIELD_UNIFORM_1D<bool>& mask) {
        for (int i = mask.i_start; i <= mask.i_end; i++) {
            if (3 <= i && i <= mask.i_end - 3) continue;
            mask(i) = true;
        }
    }
    // RemoveResidue over the finest grid (AWCM_1D.h:81-84, AWCM_MANAGER_1D.h:146-157)
    void PostProcess(FIELD_UNIFORM_1D<double>& q) {
        ITERATION_GRID_INT_1D(mra.grid_finest, 0) {
            if (mask0(i) == false) q(i) = 0;
        }
    }

    // AWCM_1D::PrepareEuler (AWCM_1D.h:86-120) on the GPU.  mask0 (the last
    // step's active set) and the field go down; the field, mask_temp,
    // mask0..mask3 and level0..level3 come back.  Throws std::runtime_error
    // carrying the AWCM status on a device error or a non-finite state.
    void PrepareEuler(FIELD_UNIFORM_1D<VECTOR_3D<double>>& q, const double& tol = 1e-6) {
        using awcm_api_detail::check;
        if (q.ghost_width != 3 || q.i_start != 0 || q.i_res != mra.Nx)
            throw std::invalid_argument("AWCM_1D::PrepareEuler: the field must span 0..Nx-1 with ghost width 3");
        if (!dev.h) check(awcm_create_1d(mra.num_levels, mra.Nx_coarse, 1.4, q.base_grid.dx, &dev.h), "awcm_create_1d");
        const size_t NP = size_t(mra.Nx) + 6;
        awcm_api_detail::push_mask0(dev.h, mask0, NP);
        check(awcm_upload_q(dev.h, q.arr.val[0].values), "awcm_upload_q");
        check(awcm_prepare_euler(dev.h, tol), "awcm_prepare_euler");
        check(awcm_download_q(dev.h, q.arr.val[0].values), "awcm_download_q");
        FIELD_UNIFORM_1D<bool>* const ms[5] = {&mask_temp, &mask0, &mask1, &mask2, &mask3};
        FIELD_UNIFORM_1D<int>* const ls[4] = {&level0, &level1, &level2, &level3};
        awcm_api_detail::fetch_masks(dev.h, NP, size_t(mra.Nx), ms, ls);
    }
    // RemoveResidue over the finest grid (AWCM_1D.h:122-125,
    // AWCM_MANAGER_1D.h:146-157)
    void PostProcessEuler(FIELD_UNIFORM_1D<VECTOR_3D<double>>& q) {
        ITERATION_GRID_INT_1D(mra.grid_finest, 0) {
            if (mask0(i) == false) q(i)[0] = q(i)[1] = q(i)[2] = 0;
        }
    }
    void Differentiate(const double& ll, const double& l, const double& c, const double& r, const double& rr,
                       double& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (rr - ll);
    }
    void DifferentiateAtLeftBoundary(const double& l, const double& c, const double& r, const double& rr, double& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (rr - l);
    }
    void DifferentiateAtRightBoundary(const double& ll, const double& l, const double& c, const double& r, double& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (r - ll);
    }
    void DifferentiateEuler(const VECTOR_3D<double>& ll, const VECTOR_3D<double>& l, const VECTOR_3D<double>& c,
                            const VECTOR_3D<double>& r, const VECTOR_3D<double>& rr, VECTOR_3D<double>& value) {
        (void)c;
        const EULER_EQUATION_1D e(1.4);
        VECTOR_3D<double> f_ll, f_l, f_r, f_rr;
        e.ComputeFu(ll, f_ll);
        e.ComputeFu(l, f_l);
        e.ComputeFu(r, f_r);
        e.ComputeFu(rr, f_rr);
        for (int p = 0; p < 3; p++) value[p] = d1 * (f_r[p] - f_l[p]) + d2 * (f_rr[p] - f_ll[p]);
    }
    void DifferentiateEulerAtLeftBoundary(const VECTOR_3D<double>& l, const VECTOR_3D<double>& c,
                                          const VECTOR_3D<double>& r, const VECTOR_3D<double>& rr,
                                          VECTOR_3D<double>& value) {
        const EULER_EQUATION_1D e(1.4);
        VECTOR_3D<double> f_l, f_c, f_r, f_rr;
        e.ComputeFu(l, f_l);
        e.ComputeFu(c, f_c);
        e.ComputeFu(r, f_r);
        e.ComputeFu(rr, f_rr);
        value = d01 * f_l + d11 * f_c + d21 * f_r + d31 * f_rr;
    }
    // the mirrored one-sided stencil keeps the left one's sign, as in
    // AWCM_1D.h:169-172
    void DifferentiateEulerAtRightBoundary(const VECTOR_3D<double>& ll, const VECTOR_3D<double>& l,
                                           const VECTOR_3D<double>& c, const VECTOR_3D<double>& r,
                                           VECTOR_3D<double>& value) {
        DifferentiateEulerAtLeftBoundary(r, c, l, ll, value);
    }
    void DifferentiateEulerAtLeftBoundary(const VECTOR_3D<double>& ll, const VECTOR_3D<double>& l,
                                          const VECTOR_3D<double>& c, const VECTOR_3D<double>& r,
                                          const VECTOR_3D<double>& rr, VECTOR_3D<double>& value) {
        const EULER_EQUATION_1D e(1.4);
        VECTOR_3D<double> f_ll, f_l, f_c, f_r, f_rr;
        e.ComputeFu(ll, f_ll);
        e.ComputeFu(l, f_l);
        e.ComputeFu(c, f_c);
        e.ComputeFu(r, f_r);
        e.ComputeFu(rr, f_rr);
        value = d02 * f_ll + d12 * f_l + d22 * f_c + d32 * f_r + d2 * f_rr;
    }
    void DifferentiateEulerAtRightBoundary(const VECTOR_3D<double>& ll, const VECTOR_3D<double>& l,
                                           const VECTOR_3D<double>& c, const VECTOR_3D<double>& r,
                                           const VECTOR_3D<double>& rr, VECTOR_3D<double>& value) {
        DifferentiateEulerAtLeftBoundary(rr, r, c, l, ll, value);
    }
};

namespace awcm_api_detail {
// The device solver of a 2-D field: validates the field's extent and creates
// the solver on first use (dx and the x of column 0 come from the field's grid).
template <class T>
awcm_solver* open_2d(Handle& dev, const MRA_GRID_INT_2D& mra_grid, const FIELD_UNIFORM_2D<T>& q, const char* who) {
    const int Nx = mra_grid.Nx, Ny = mra_grid.Ny;
    if (q.ghost_width != 3 || q.i_start != 0 || q.j_start != 0 || q.i_res != Nx || q.j_res != Ny)
        throw std::invalid_argument(std::string(who) +
                                    ": the field must span (0..Nx-1) x (0..Ny-1) with ghost width 3");
    if (!dev.h)
        check(awcm_create_2d(mra_grid.num_levels, mra_grid.Nx_coarse, mra_grid.Ny_coarse, 1.4, q.grid.dx,
                             q.grid.x_min, AWCM_BC_TRANSMISSIVE, &dev.h),
              "awcm_create_2d");
    return dev.h;
}
// A mask argument spans the finest grid; its ghost width may be anything (the
// drivers pass ghost-3 masks and, main_2d_advect.cpp:217, a FIELD_UNIFORM_2D
// built on the grid alone, ghost width 0).  Only its finest-grid points are
// read, as in the reference.
inline void check_mask_2d(const MRA_GRID_INT_2D& mra_grid, const FIELD_UNIFORM_2D<bool>& m, const char* who) {
    if (m.i_start != 0 || m.j_start != 0 || m.i_res != mra_grid.Nx || m.j_res != mra_grid.Ny)
        throw std::invalid_argument(std::string(who) + ": masks must span (0..Nx-1) x (0..Ny-1)");
}
// the mask in the C ABI's layout: (Nx+6) x (Ny+6) bytes, ghosts it lacks zero
inline std::unique_ptr<uint8_t[]> mask_bytes(const FIELD_UNIFORM_2D<bool>& m) {
    const int Nx = m.i_res, Ny = m.j_res, g = m.ghost_width < 3 ? m.ghost_width : 3;
    const size_t P = size_t(Nx) + 6;
    std::unique_ptr<uint8_t[]> b(new uint8_t[P * (size_t(Ny) + 6)]());
    for (int j = -g; j <= Ny - 1 + g; j++)
        for (int i = -g; i <= Nx - 1 + g; i++) b[size_t(i + 3) + P * size_t(j + 3)] = m(i, j) ? 1 : 0;
    return b;
}
}  // namespace awcm_api_detail

// WAVELET_2D (Wavelet/WAVELET_2D.h:10-35): the interpolating wavelet of order
// 2N.  N and weight(i, j) = l_i(j + 1/2), the Lagrange basis on the nodes
// 0..2N-1 evaluated at the midpoints, are the reference's public members; the
// device kernels hold their own copy of the N = 2 table, the only order the
// AWCM classes use.  Of its transforms, the interpolations are exposed (the
// drivers call them directly: main_2d_euler.cpp:334, main_2d_advect.cpp:219);
// the forward and inverse transforms run inside AWCM_2D::Prepare and
// PrepareEuler on the device.  A WAVELET_2D that is AWCM_2D's member
// shares that object's device solver; a free-standing one has its own.
class WAVELET_2D {
public:
    const int N;
    ARRAY_2D<double> weight;

    WAVELET_2D(const int& N_ = 2) : N(N_) {
        if (N != 2) throw std::invalid_argument("WAVELET_2D: only N = 2 (the 4-point wavelet) is supported");
        weight.Initialize(0, -1, 2 * N, 2 * N + 1);
        for (int j = -1; j < 2 * N; j++)
            for (int i = 0; i < 2 * N; i++) {
                double l = 1.0;
                for (int k = 0; k < 2 * N; k++)
                    if (k != i) l *= double(j + 0.5 - k) / double(i - k);
                weight(i, j) = l;
            }
    }
    WAVELET_2D(const WAVELET_2D&) = delete;
    WAVELET_2D& operator=(const WAVELET_2D&) = delete;

    // WAVELET_2D::Interpolate (WAVELET_2D.h:360-473) on the GPU: the scalar
    // field at the points of CheckReconstruction(mask_to minus mask_from) not
    // in mask_from.  The field goes down and comes back.
    void Interpolate(FIELD_UNIFORM_2D<double>& q, const MRA_GRID_INT_2D& mra_grid,
                     const FIELD_UNIFORM_2D<bool>& mask_from, const FIELD_UNIFORM_2D<bool>& mask_to) const {
        using awcm_api_detail::check;
        const char* who = "WAVELET_2D::Interpolate";
        awcm_api_detail::check_mask_2d(mra_grid, mask_from, who);
        awcm_api_detail::check_mask_2d(mra_grid, mask_to, who);
        awcm_solver* h = awcm_api_detail::open_2d(*dev, mra_grid, q, who);
        check(awcm_upload_scalar(h, q.arr.val), "awcm_upload_scalar");
        check(awcm_interpolate(h, awcm_api_detail::mask_bytes(mask_from).get(),
                               awcm_api_detail::mask_bytes(mask_to).get()),
              "awcm_interpolate");
        check(awcm_download_scalar(h, q.arr.val), "awcm_download_scalar");
    }
    // WAVELET_2D::InterpolateEuler (WAVELET_2D.h:784-907) on the GPU: the same
    // for momentum and energy.
    void InterpolateEuler(FIELD_UNIFORM_2D<VECTOR_4D<double>>& q, const MRA_GRID_INT_2D& mra_grid,
                          const FIELD_UNIFORM_2D<bool>& mask_from, const FIELD_UNIFORM_2D<bool>& mask_to) const {
        using awcm_api_detail::check;
        const char* who = "WAVELET_2D::InterpolateEuler";
        awcm_api_detail::check_mask_2d(mra_grid, mask_from, who);
        awcm_api_detail::check_mask_2d(mra_grid, mask_to, who);
        awcm_solver* h = awcm_api_detail::open_2d(*dev, mra_grid, q, who);
        check(awcm_upload_q(h, q.arr.val[0].values), "awcm_upload_q");
        check(awcm_interpolate_euler(h, awcm_api_detail::mask_bytes(mask_from).get(),
                                     awcm_api_detail::mask_bytes(mask_to).get()),
              "awcm_interpolate_euler");
        check(awcm_download_q(h, q.arr.val[0].values), "awcm_download_q");
    }
    // WAVELET_2D::InterpolateAllGridEuler (WAVELET_2D.h:912-1044) on the GPU:
    // every finest-grid point not in mask_from gets its prediction from the
    // coarser levels, all four components.  The field goes down and comes back.
    void InterpolateAllGridEuler(FIELD_UNIFORM_2D<VECTOR_4D<double>>& q, const MRA_GRID_INT_2D& mra_grid,
                                 const FIELD_UNIFORM_2D<bool>& mask_from) const {
        using awcm_api_detail::check;
        const char* who = "WAVELET_2D::InterpolateAllGridEuler";
        awcm_api_detail::check_mask_2d(mra_grid, mask_from, who);
        awcm_solver* h = awcm_api_detail::open_2d(*dev, mra_grid, q, who);
        check(awcm_upload_q(h, q.arr.val[0].values), "awcm_upload_q");
        check(awcm_interpolate_all_grid(h, awcm_api_detail::mask_bytes(mask_from).get()), "awcm_interpolate_all_grid");
        check(awcm_download_q(h, q.arr.val[0].values), "awcm_download_q");
    }

private:
    friend class AWCM_2D;
    mutable awcm_api_detail::Handle own;
    awcm_api_detail::Handle* dev = &own;
};

class AWCM_2D {
public:
    WAVELET_2D wavelet;
    MRA_GRID_INT_2D& mra_grid;
    FIELD_UNIFORM_2D<bool> mask_temp, mask0, mask1, mask2, mask3;
    FIELD_UNIFORM_2D<int> level0, level1, level2, level3;
    const double d1, d2;
    double d00, d01, d02;
    double d10, d11, d12, d13;
    double d20, d21, d22, d23, d24;
    double d30, d31, d32, d33, d34, d35;
    awcm_api_detail::Handle dev;

    // gamma is accepted and unused, as in the reference (AWCM_2D.h:24)
    AWCM_2D(MRA_GRID_INT_2D& mra_grid_, const double& gamma = 1.4)
        : mra_grid(mra_grid_), d1(2.0 / 3.0), d2(-1.0 / 12.0) {
        (void)gamma;
        wavelet.dev = &dev;
        const int Nx = mra_grid.Nx, Ny = mra_grid.Ny;
        FIELD_UNIFORM_2D<bool>* ms[5] = {&mask_temp, &mask0, &mask1, &mask2, &mask3};
        FIELD_UNIFORM_2D<int>* ls[4] = {&level0, &level1, &level2, &level3};
        for (auto* m : ms) m->Initialize(0, 0, Nx, Ny, 0, 0, Nx, Ny, 3);
        for (auto* l : ls) {
            l->Initialize(0, 0, Nx, Ny, 0, 0, Nx, Ny);
            l->AssignAllValues(mra_grid.num_levels - 1);
        }
        mask0.AssignAllValues(true);
        d00 = -11.0 / 6.0; d01 = -1.0 / 3.0; d02 = 1.0 / 12.0;
        d10 = 3.0; d11 = -1.0 / 2.0; d12 = -2.0 / 3.0; d13 = 1.0 / 12.0;
        d20 = -3.0 / 2.0; d21 = 1.0; d22 = 0; d23 = -2.0 / 3.0; d24 = 1.0 / 12.0;
        d30 = 1.0 / 3.0; d31 = -1.0 / 6.0; d32 = 2.0 / 3.0; d33 = 0; d34 = -2.0 / 3.0; d35 = 1.0 / 12.0;
    }

    // AWCM_2D::Prepare (AWCM_2D.h:52-79) on the GPU: the scalar path.  mask0
    // and the field go down; the field, mask0..mask3 and level0..level3 come
    // back (mask_temp is not the scalar path's and stays as it was).  Throws
    // std::runtime_error carrying the AWCM status on a device error.
    void Prepare(FIELD_UNIFORM_2D<double>& q, const double& tol = 1e-6) {
        using awcm_api_detail::check;
        const int Nx = mra_grid.Nx, Ny = mra_grid.Ny;
        awcm_api_detail::open_2d(dev, mra_grid, q, "AWCM_2D::Prepare");
        const size_t NP = size_t(Nx + 6) * size_t(Ny + 6);
        awcm_api_detail::push_mask0(dev.h, mask0, NP);
        check(awcm_upload_scalar(dev.h, q.arr.val), "awcm_upload_scalar");
        check(awcm_prepare(dev.h, tol), "awcm_prepare");
        check(awcm_download_scalar(dev.h, q.arr.val), "awcm_download_scalar");
        FIELD_UNIFORM_2D<bool> keep(mask_temp);
        FIELD_UNIFORM_2D<bool>* const ms[5] = {&mask_temp, &mask0, &mask1, &mask2, &mask3};
        FIELD_UNIFORM_2D<int>* const ls[4] = {&level0, &level1, &level2, &level3};
        awcm_api_detail::fetch_masks(dev.h, NP, size_t(Nx) * size_t(Ny), ms, ls);
        mask_temp = keep;
    }
    // RemoveResidue over the finest grid minus its outer ring (AWCM_2D.h:81-84,
    // AWCM_MANAGER_2D.h:257-266)
    void PostProcess(FIELD_UNIFORM_2D<double>& q) {
        ITERATION_GRID_INT_2D(mra_grid.grid_finest, -1) {
            if (mask0(i, j) == false) q(i, j) = 0.0;
        }
    }
    // AWCM_2D.h:85-93: the 2-wide band next to the walls is always active
    void BandingBoundary(FIELD_UNIFORM_2D<bool>& mask) {
        for (int j = mask.j_start; j <= mask.j_end; j++)
            for (int i = mask.i_start; i <= mask.i_end; i++) {
                if (2 <= i && i <= mask.i_end - 2 && j >= 2 && j <= mask.j_end - 2) continue;
                mask(i, j) = true;
            }
    }
    // the scalar forms of AWCM_2D.h:95-110
    void Differentiate(const double& ll, const double& l, const double& c, const double& r, const double& rr,
                       double& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (rr - ll);
    }
    void DifferentiateAtLeftBoundary(const double& l, const double& c, const double& r, const double& rr,
                                     double& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (rr - l);
    }
    void DifferentiateAtRightBoundary(const double& ll, const double& l, const double& c, const double& r,
                                      double& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (r - ll);
    }

    // AWCM_2D::PrepareEuler (AWCM_2D.h:114-147) on the GPU.  mask0 and the
    // field go down; the field, mask_temp, mask0..mask3 and level0..level3
    // come back.  Throws std::runtime_error carrying the AWCM status on a
    // device error or a non-finite state.
    void PrepareEuler(FIELD_UNIFORM_2D<VECTOR_4D<double>>& q, const double& tol = 1e-6) {
        using awcm_api_detail::check;
        const int Nx = mra_grid.Nx, Ny = mra_grid.Ny;
        awcm_api_detail::open_2d(dev, mra_grid, q, "AWCM_2D::PrepareEuler");
        const size_t NP = size_t(Nx + 6) * size_t(Ny + 6);
        awcm_api_detail::push_mask0(dev.h, mask0, NP);
        check(awcm_upload_q(dev.h, q.arr.val[0].values), "awcm_upload_q");
        check(awcm_prepare_euler(dev.h, tol), "awcm_prepare_euler");
        check(awcm_download_q(dev.h, q.arr.val[0].values), "awcm_download_q");
        FIELD_UNIFORM_2D<bool>* const ms[5] = {&mask_temp, &mask0, &mask1, &mask2, &mask3};
        FIELD_UNIFORM_2D<int>* const ls[4] = {&level0, &level1, &level2, &level3};
        awcm_api_detail::fetch_masks(dev.h, NP, size_t(Nx) * size_t(Ny), ms, ls);
    }
    // RemoveResidueEuler over the finest grid minus its outer ring
    // (AWCM_2D.h:149-152, AWCM_MANAGER_2D.h:267-280)
    void PostProcessEuler(FIELD_UNIFORM_2D<VECTOR_4D<double>>& q) {
        ITERATION_GRID_INT_2D(mra_grid.grid_finest, -1) {
            if (mask0(i, j) == false) q(i, j)[0] = q(i, j)[1] = q(i, j)[2] = q(i, j)[3] = 0.0;
        }
    }
    void Differentiate(const VECTOR_4D<double>& ll, const VECTOR_4D<double>& l, const VECTOR_4D<double>& c,
                       const VECTOR_4D<double>& r, const VECTOR_4D<double>& rr, VECTOR_4D<double>& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (rr - ll);
    }
    // the one-sided forms of AWCM_2D.h:166-187 as written there
    void DifferentiateAtLeftBoundary(const VECTOR_4D<double>& l, const VECTOR_4D<double>& c, const VECTOR_4D<double>& r,
                                     const VECTOR_4D<double>& rr, VECTOR_4D<double>& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (rr - l);
    }
    void DifferentiateAtRightBoundary(const VECTOR_4D<double>& ll, const VECTOR_4D<double>& l,
                                      const VECTOR_4D<double>& c, const VECTOR_4D<double>& r,
                                      VECTOR_4D<double>& value) {
        (void)c;
        value = d1 * (r - l) + d2 * (r - ll);
    }
};

/* ---- problems and their boundary managers (Problems/1D, Problems/2D) ------ */
namespace awcm_api_detail {
// quadrant states (rho, u, v, p) for NE, NW, SW, SE (PROBLEM_QUADRANTS.h:37-73)
inline const double* quadrant_states() {
    static const double s[16] = {1.5, 0, 0, 1.5,
                                 0.532258064516129, 1.206045378311055, 0, 0.3,
                                 0.137992831541219, 1.206045378311055, 1.206045378311055, 0.029032258064516,
                                 0.532258064516129, 0, 1.206045378311055, 0.3};
    return s;
}
// the cell-centred 1-D problem grids: res+1 faces on [xl, xu], res centres
inline void cell_grids(int res, double xl, double xu, GRID_UNIFORM_1D& grid, GRID_UNIFORM_1D& grid_x) {
    grid_x.Initialize(0, res + 1, xl, xu);
    const double dx = grid_x.dx;
    grid.Initialize(0, res, xl + 0.5 * dx, xu - 0.5 * dx);
}
// the node-centred 2-D problem grids and their x-, y- and xy-staggered twins
inline void node_grids(int nx, int ny, double xl, double yl, double xu, double yu, GRID_UNIFORM_2D& grid,
                       GRID_UNIFORM_2D& grid_x, GRID_UNIFORM_2D& grid_y, GRID_UNIFORM_2D& grid_xy) {
    grid.Initialize(0, 0, nx, ny, xl, yl, xu, yu);
    const double dx = grid.dx, dy = grid.dy;
    grid_x.Initialize(0, 0, nx + 1, ny, xl - 0.5 * dx, yl, xu + 0.5 * dx, yu);
    grid_y.Initialize(0, 0, nx, ny + 1, xl, yl - 0.5 * dy, xu, yu + 0.5 * dy);
    grid_xy.Initialize(0, 0, nx + 1, ny + 1, xl - 0.5 * dx, yl - 0.5 * dy, xu + 0.5 * dx, yu + 0.5 * dy);
}
}  // namespace awcm_api_detail

// PROBLEM_SHOCKTUBE.h:14-54: Sod on [0, 10], left state where x <= 5,
// t_final = 2
class SHOCKTUBE {
public:
    static void Initialize(const int& resolution, GRID_UNIFORM_1D& grid, GRID_UNIFORM_1D& grid_x,
                           FIELD_UNIFORM_1D<VECTOR_3D<double>>& Q, double& t_final, double& gamma) {
        t_final = 2;
        gamma = 1.4;
        awcm_api_detail::cell_grids(resolution, 0.0, 10.0, grid, grid_x);
        Q.Initialize(grid, 3);
        const double left[3] = {1, 0, 1}, right[3] = {0.125, 0, 0.1};
        awcm_api_detail::check(awcm_ic_riemann_1d(resolution, 0.0, 10.0, 5.0, left, right, gamma, Q.arr.val[0].values,
                                                  nullptr, nullptr),
                               "awcm_ic_riemann_1d");
    }
};

// PROBLEM_ENTROPY_INTERACTING.h:14-49: Shu-Osher on [-5, 5], shocked state
// where x < -4, t_final = 1.8
class ENTROPY_INTERACTING {
public:
    static void Initialize(const int& resolution, GRID_UNIFORM_1D& grid, GRID_UNIFORM_1D& grid_x,
                           FIELD_UNIFORM_1D<VECTOR_3D<double>>& Q, double& t_final, double& gamma) {
        t_final = 1.8;
        gamma = 1.4;
        awcm_api_detail::cell_grids(resolution, -5.0, 5.0, grid, grid_x);
        Q.Initialize(grid, 3);
        const double left[3] = {3.857143, 2.629369, 10.33333};
        awcm_api_detail::check(awcm_ic_shu_osher_1d(resolution, -5.0, 5.0, -4.0, left, 0.2, gamma, Q.arr.val[0].values,
                                                    nullptr, nullptr),
                               "awcm_ic_shu_osher_1d");
    }
};

// zero-order extrapolation into the ghosts at both ends
// (PROBLEM_SHOCKTUBE.h:57-107, PROBLEM_ENTROPY_INTERACTING.h:52-102).  The
// reference names the Sod one BOUNDARY_MANAGER, as it names the periodic
// scalar one of PROBLEM_BOX_ADVECTION.h; each is declared by its problem's
// header (include/compat/Problems/...), as there, so the two names never meet.
namespace awcm_api_detail {
class Extrap1D {
public:
    int i_start, i_start_g, i_end, i_end_g, ghost_width;
    FIELD_UNIFORM_1D<VECTOR_3D<double>>& Q;
    Extrap1D(FIELD_UNIFORM_1D<VECTOR_3D<double>>& Q_)
        : i_start(Q_.i_start), i_start_g(Q_.i_start_g), i_end(Q_.i_end), i_end_g(Q_.i_end_g),
          ghost_width(Q_.ghost_width), Q(Q_) {}
    void SetLeftBoundaryCondition() {
        for (int i = i_start - 1; i >= i_start_g; i--) Q(i) = Q(i_start);
    }
    void SetRightBoundaryCondition() {
        for (int i = i_end + 1; i <= i_end_g; i++) Q(i) = Q(i_end);
    }
    void SetBoundaryCondition() {
        SetLeftBoundaryCondition();
        SetRightBoundaryCondition();
    }
};
}  // namespace awcm_api_detail
// the reference's class name, spelling included
class BOUNDARY_MANAGER_ENTROPY_INTERATING : public awcm_api_detail::Extrap1D {
public:
    BOUNDARY_MANAGER_ENTROPY_INTERATING(FIELD_UNIFORM_1D<VECTOR_3D<double>>& Q_) : Extrap1D(Q_) {}
};

// PROBLEM_QUADRANTS.h:15-81: [0, 1]^2, quadrants split at 0.8, t_final = 0.8
class QUADRANTS {
public:
    static void Initialize(const int& resolution, GRID_UNIFORM_2D& grid, GRID_UNIFORM_2D& grid_x,
                           GRID_UNIFORM_2D& grid_y, GRID_UNIFORM_2D& grid_xy, FIELD_UNIFORM_2D<VECTOR_4D<double>>& Q,
                           double& t_final, double& gamma) {
        gamma = 1.4;
        t_final = 0.8;
        awcm_api_detail::node_grids(resolution, resolution, 0.0, 0.0, 1.0, 1.0, grid, grid_x, grid_y, grid_xy);
        Q.Initialize(grid, 3);
        awcm_api_detail::check(awcm_ic_quadrants_2d(resolution, 0.8, awcm_api_detail::quadrant_states(), gamma,
                                                    Q.arr.val[0].values, nullptr),
                               "awcm_ic_quadrants_2d");
    }
};

// PROBLEM_DOUBLEMACH.h:13-63: [0, 4] x [0, 1], the Mach-10 shock leaving
// x = 1/6 at 60 degrees, pre-shock density 1.5 as that file sets it,
// t_final = 0.2
class DOUBLEMACH {
public:
    static void Initialize(const int& resolution, GRID_UNIFORM_2D& grid, GRID_UNIFORM_2D& grid_x,
                           GRID_UNIFORM_2D& grid_y, GRID_UNIFORM_2D& grid_xy, FIELD_UNIFORM_2D<VECTOR_4D<double>>& Q,
                           double& t_final, double& gamma) {
        gamma = 1.4;
        t_final = 0.2;
        awcm_api_detail::node_grids((resolution - 1) * 4 + 1, resolution, 0.0, 0.0, 4.0, 1.0, grid, grid_x, grid_y,
                                    grid_xy);
        Q.Initialize(grid, 3);
        awcm_api_detail::check(awcm_ic_double_mach_2d(resolution, 1.5, gamma, Q.arr.val[0].values, nullptr),
                               "awcm_ic_double_mach_2d");
    }
};

// PROBLEM_BOX_ADVECTION.h:5-56 (the scalar path's problem): a Gaussian
// exp(-50 (x-1/2)^2 - 50 (y-1/2)^2) on the cell centres of [0, 1]^2, ghost
// cells included, t_final = 2.  The cell-centre grid's y_max is written with
// x_upper there; both are 1.
class BOX_ADVECTION {
public:
    static void Initialize(const int& resolution_x, GRID_UNIFORM_2D& grid, GRID_UNIFORM_2D& grid_x,
                           GRID_UNIFORM_2D& grid_y, GRID_UNIFORM_2D& grid_xy, FIELD_UNIFORM_2D<double>& Q,
                           double& t_final) {
        const double x_lower = 0, x_upper = 1, y_lower = 0, y_upper = 1;
        t_final = 2;
        grid_xy.Initialize(0, 0, resolution_x + 1, resolution_x + 1, x_lower, y_lower, x_upper, y_upper);
        const double dx = grid_xy.dx, dy = grid_xy.dy;
        grid_x.Initialize(0, 0, resolution_x + 1, resolution_x, x_lower, y_lower + 0.5 * dy, x_upper, y_upper - 0.5 * dy);
        grid_y.Initialize(0, 0, resolution_x, resolution_x + 1, x_lower + 0.5 * dx, y_lower, x_upper - 0.5 * dx, y_upper);
        grid.Initialize(0, 0, resolution_x, resolution_x, x_lower + 0.5 * dx, y_lower + 0.5 * dy, x_upper - 0.5 * dx,
                        x_upper - 0.5 * dy);
        Q.Initialize(grid, 3);
        for (int j = grid.j_start - 3; j <= grid.j_end + 3; j++)
            for (int i = grid.i_start - 3; i <= grid.i_end + 3; i++) {
                const VECTOR_2D<double> c = grid(i, j);
                Q(i, j) = std::exp(-50 * SQR(c.x - 0.5) - 50 * SQR(c.y - 0.5));
            }
    }
};

namespace awcm_api_detail {
struct Bounds2 {
    int i_start, i_start_g, i_end, i_end_g, j_start, j_start_g, j_end, j_end_g, ghost_width;
    template <class T>
    explicit Bounds2(const FIELD_UNIFORM_2D<T>& Q)
        : i_start(Q.i_start), i_start_g(Q.i_start_g), i_end(Q.i_end), i_end_g(Q.i_end_g), j_start(Q.j_start),
          j_start_g(Q.j_start_g), j_end(Q.j_end), j_end_g(Q.j_end_g), ghost_width(Q.ghost_width) {}
};
}  // namespace awcm_api_detail

// PROBLEM_QUADRANTS.h:85-203: each side extrapolates its edge row or column;
// the corner blocks are filled only by SetDiagonalBoundaryCondition, which
// SetBoundaryCondition does not call.
class BOUNDARY_MANAGER_QUAD : public awcm_api_detail::Bounds2 {
public:
    FIELD_UNIFORM_2D<VECTOR_4D<double>>& Q;
    BOUNDARY_MANAGER_QUAD(FIELD_UNIFORM_2D<VECTOR_4D<double>>& Q_) : Bounds2(Q_), Q(Q_) {}
    void SetLeftBoundaryCondition() {
#pragma omp parallel for
        for (int j = j_start; j <= j_end; j++)
            for (int i = i_start - 1; i >= i_start_g; i--) Q(i, j) = Q(i_start, j);
    }
    void SetRightBoundaryCondition() {
#pragma omp parallel for
        for (int j = j_start; j <= j_end; j++)
            for (int i = i_end + 1; i <= i_end_g; i++) Q(i, j) = Q(i_end, j);
    }
    void SetBottomBoundaryCondition() {
#pragma omp parallel for
        for (int i = i_start; i <= i_end; i++)
            for (int j = j_start - 1; j >= j_start_g; j--) Q(i, j) = Q(i, j_start);
    }
    void SetTopBoundaryCondition() {
#pragma omp parallel for
        for (int i = i_start; i <= i_end; i++)
            for (int j = j_end + 1; j <= j_end_g; j++) Q(i, j) = Q(i, j_end);
    }
    void SetDiagonalBoundaryCondition() {
        for (int i = i_start_g; i < i_start; i++) {
            for (int j = j_start_g; j < j_start; j++) Q(i, j) = Q(i_start, j_start);
            for (int j = j_end + 1; j <= j_end_g; j++) Q(i, j) = Q(i_start, j_end);
        }
        for (int i = i_end + 1; i <= i_end_g; i++) {
            for (int j = j_start_g; j < j_start; j++) Q(i, j) = Q(i_end, j_start);
            for (int j = j_end + 1; j <= j_end_g; j++) Q(i, j) = Q(i_end, j_end);
        }
    }
    void SetBoundaryCondition() {
        SetLeftBoundaryCondition();
        SetRightBoundaryCondition();
        SetBottomBoundaryCondition();
        SetTopBoundaryCondition();
    }
};

// PROBLEM_DOUBLEMACH.h:66-245: post-shock inflow on the left, outflow on the
// right; at the bottom the post-shock state up to x = 1/6 and a reflecting
// wall beyond it (every ghost row mirrors row j_start); at the top the post-
// or pre-shock state on either side of the moving shock
// x = 1/6 + (1+20t)/sqrt(3).  SetBoundaryCondition leaves the corner blocks.
class BOUNDARY_MANAGER_DMACH : public awcm_api_detail::Bounds2 {
public:
    const IDEAL_GAS_2D EOS;
    FIELD_UNIFORM_2D<VECTOR_4D<double>>& Q;
    BOUNDARY_MANAGER_DMACH(FIELD_UNIFORM_2D<VECTOR_4D<double>>& Q_) : Bounds2(Q_), EOS(1.4), Q(Q_) {}

    void SetLeftBoundaryCondition() {
        const double rho = 8, u = 8.25 * std::sqrt(3.0) * 0.5, v = -8.25 * 0.5, p = 116.5;
        for (int j = j_start; j <= j_end; j++)
            for (int i = i_start - 1; i >= i_start_g; i--) set(i, j, rho, rho * u, rho * v, EOS.E(rho, u, v, p));
    }
    void SetRightBoundaryCondition() {
        for (int j = j_start; j <= j_end; j++)
            for (int i = i_end + 1; i <= i_end_g; i++) Q(i, j) = Q(i_end, j);
    }
    void SetBottomBoundaryCondition() {
        const GRID_UNIFORM_2D& grid(Q.grid);
        const double sqrt3 = std::sqrt(3.0);
        for (int i = i_start; i <= i_end; i++)
            for (int j = j_start - 1; j >= j_start_g; j--) {
                if (grid(i, j).x > 1.0 / 6.0) {
                    const VECTOR_4D<double> w = Q(i, j_start);
                    set(i, j, w[0], w[1], -w[2], w[3]);
                } else {
                    post_shock(i, j, sqrt3);
                }
            }
    }
    void SetTopBoundaryCondition(const double& t) {
        const GRID_UNIFORM_2D& grid(Q.grid);
        const double sqrt3 = std::sqrt(3.0);
        for (int i = i_start; i <= i_end; i++)
            for (int j = j_end + 1; j <= j_end_g; j++) {
                if (grid(i, j).x < 1.0 / 6.0 + (1 + 20 * t) / sqrt3) post_shock(i, j, sqrt3);
                else set(i, j, 1.4, 1.4 * 0, 1.4 * 0, EOS.E(1.4, 0, 0, 1));
            }
    }
    // corner blocks copied across the domain (PROBLEM_DOUBLEMACH.h:178-208);
    // not part of SetBoundaryCondition
    void SetDiagonalBoundaryCondition() {
        int c1 = 0;
        for (int i = i_start - 1; i >= i_start_g; i--, c1++) {
            int c2 = 0;
            for (int j = j_start - 1; j >= j_start_g; j--, c2++) Q(i, j) = Q(i_end - c1, j_end - c2);
            c2 = 0;
            for (int j = j_end + 1; j <= j_end_g; j++, c2++) Q(i, j) = Q(i_end - c1, j_start + c2);
        }
        c1 = 0;
        for (int i = i_end + 1; i <= i_end_g; i++, c1++) {
            int c2 = 0;
            for (int j = j_start - 1; j >= j_start_g; j--, c2++) Q(i, j) = Q(i_start + c1, j_end - c2);
            for (int j = j_end + 1; j <= j_end_g; j++, c2++) Q(i, j) = Q(i_start + c1, j_start + c2);
        }
    }
    void SetBoundaryCondition(const double& t) {
        SetLeftBoundaryCondition();
        SetRightBoundaryCondition();
        SetBottomBoundaryCondition();
        SetTopBoundaryCondition(t);
    }

private:
    void set(int i, int j, double a, double b, double c, double d) {
        VECTOR_4D<double>& q = Q(i, j);
        q[0] = a; q[1] = b; q[2] = c; q[3] = d;
    }
    void post_shock(int i, int j, double sqrt3) {
        set(i, j, 8, 8 * 8.25 * sqrt3 * 0.5, -8 * 8.25 * 0.5, EOS.E(8, 8.25 * sqrt3 * 0.5, -8.25 * 0.5, 116.5));
    }
};

// PROBLEM_BOX_ADVECTION.h:59-197: periodic ghost fills of the scalar field
// (declared as BOUNDARY_MANAGER by include/compat/Problems/2D/
// PROBLEM_BOX_ADVECTION.h).  SetBoundaryCondition does the four sides; the
// corner blocks are filled only by SetDiagonalBoundaryCondition, whose second
// pass on the right side carries its j counter over from the first, as the
// reference's does.
namespace awcm_api_detail {
class Periodic2D : public Bounds2 {
public:
    FIELD_UNIFORM_2D<double>& Q;
    Periodic2D(FIELD_UNIFORM_2D<double>& Q_) : Bounds2(Q_), Q(Q_) {}
    void Initialize(const int i_start_, const int i_start_g_, const int i_end_, const int i_end_g_, const int j_start_,
                    const int j_start_g_, const int j_end_, const int j_end_g_, const int ghost_width_) {
        i_start = i_start_; i_start_g = i_start_g_; i_end = i_end_; i_end_g = i_end_g_;
        j_start = j_start_; j_start_g = j_start_g_; j_end = j_end_; j_end_g = j_end_g_;
        ghost_width = ghost_width_;
    }
    void SetLeftBoundaryCondition() {
        for (int j = j_start; j <= j_end; j++) {
            int c = 0;
            for (int i = i_start - 1; i >= i_start_g; i--, c++) Q(i, j) = Q(i_end - c, j);
        }
    }
    void SetRightBoundaryCondition() {
        for (int j = j_start; j <= j_end; j++) {
            int c = 0;
            for (int i = i_end + 1; i <= i_end_g; i++, c++) Q(i, j) = Q(i_start + c, j);
        }
    }
    void SetBottomBoundaryCondition() {
        for (int i = i_start; i <= i_end; i++) {
            int c = 0;
            for (int j = j_start - 1; j >= j_start_g; j--, c++) Q(i, j) = Q(i, j_end - c);
        }
    }
    void SetTopBoundaryCondition() {
        for (int i = i_start; i <= i_end; i++) {
            int c = 0;
            for (int j = j_end + 1; j <= j_end_g; j++, c++) Q(i, j) = Q(i, j_start + c);
        }
    }
    void SetDiagonalBoundaryCondition() {
        int c1 = 0;
        for (int i = i_start - 1; i >= i_start_g; i--, c1++) {
            int c2 = 0;
            for (int j = j_start - 1; j >= j_start_g; j--, c2++) Q(i, j) = Q(i_end - c1, j_end - c2);
            c2 = 0;
            for (int j = j_end + 1; j <= j_end_g; j++, c2++) Q(i, j) = Q(i_end - c1, j_start + c2);
        }
        c1 = 0;
        for (int i = i_end + 1; i <= i_end_g; i++, c1++) {
            int c2 = 0;
            for (int j = j_start - 1; j >= j_start_g; j--, c2++) Q(i, j) = Q(i_start + c1, j_end - c2);
            for (int j = j_end + 1; j <= j_end_g; j++, c2++) Q(i, j) = Q(i_start + c1, j_start + c2);
        }
    }
    void SetBoundaryCondition() {
        SetLeftBoundaryCondition();
        SetRightBoundaryCondition();
        SetBottomBoundaryCondition();
        SetTopBoundaryCondition();
    }
};
}  // namespace awcm_api_detail
