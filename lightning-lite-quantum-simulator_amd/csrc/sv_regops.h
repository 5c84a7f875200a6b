// sv_regops.h -- register-tile arithmetic shared by the compiled-in pass interpreter
// (k_pass_reg) and the source text of generated pass kernels (sv_jit.hip).
//
// LLSV_REGOPS_SOURCE is expanded as code in sv_program.hip and stringized into the prelude of
// every generated kernel, so both paths run one definition and produce identical bits.  Each
// helper spells out its fused multiply-adds, so the rounding sequence does not depend on the
// compiler's contraction choices.  Needs v2d (2 x double ext vector) in scope.
#pragma once

// r_swz: XOR swizzle of an LDS slot (bits 4..7 of a local index are folded into bits 0..3, so
//   a wavefront writing along bits 4..7 spreads over the banks); XOR-linear on disjoint bits.
// r_ins0: inserts a zero bit at position pos.
// r_pu: (x, y) <- (m00 x + m01 y, m10 x + m11 y), m row-major interleaved re/im.  Every output
//   component is finished by the FMA that last reads the input component it replaces.
// r_rx: (x, y) <- (p x + i q y, i q x + p y), the form of RX and its inverse (p, q real).
// r_ry: (x, y) <- (p x - q y, q x + p y), the form of RY and its inverse (p, q real).
// r_sc: v <- (dr + i di) v.
// r_ng: flips the sign of both components when sgn = 0x80000000 (a sign-bit XOR, as -v);
//   the components go through scalar locals (a bit cast of a vector element expression such
//   as v.y reads element 0 under this compiler).
// r_phase: (sr + i si) <- (sr + i si)(dr + i di).
// r_shy / r_shx: the rotation [[cos f, -sin f], [sin f, cos f]] (RY form) or
//   [[cos f, -i sin f], [-i sin f, cos f]] (RX form) of an angle |f| <= pi/2 as three shears,
//   x += a y, y += b x, x += a y (the RX form with i a, i b), with a = -tan(f/2) and b = sin f
//   (RY) or -sin f (RX): six FMAs per pair, in place.
// r_nu0 / r_nu1: a one-qubit matrix with its row scales taken out (normalised form), one FMA
//   per product: r_nu0 (x, y) <- (x + b y, c x + y), r_nu1 (x, y) <- (b x + y, x + c y), with
//   b = m[0] + i m[1], c = m[2] + i m[3].
#define LLSV_REGOPS_SOURCE                                                                              \
    __device__ __forceinline__ unsigned int r_swz(unsigned int l) { return l ^ ((l >> 4) & 15u); }      \
    __device__ __forceinline__ unsigned int r_ins0(unsigned int x, int pos) {                          \
        return ((x >> pos) << (pos + 1)) | (x & ((1u << pos) - 1u));                                    \
    }                                                                                                   \
    __device__ __forceinline__ void r_pu(v2d& x, v2d& y, const double* m) {                             \
        const double px_re = __builtin_fma(-m[3], y.y, __builtin_fma(m[2], y.x, -m[1] * x.y));          \
        const double px_im = __builtin_fma(m[3], y.x, __builtin_fma(m[2], y.y, m[1] * x.x));            \
        const double py_re = __builtin_fma(-m[7], y.y, __builtin_fma(-m[5], x.y, m[4] * x.x));          \
        const double py_im = __builtin_fma(m[7], y.x, __builtin_fma(m[5], x.x, m[4] * x.y));            \
        x.x = __builtin_fma(m[0], x.x, px_re);                                                          \
        x.y = __builtin_fma(m[0], x.y, px_im);                                                          \
        y.x = __builtin_fma(m[6], y.x, py_re);                                                          \
        y.y = __builtin_fma(m[6], y.y, py_im);                                                          \
    }                                                                                                   \
    __device__ __forceinline__ void r_rx(v2d& x, v2d& y, double p, double q) {                          \
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;                                            \
        x.x = __builtin_fma(p, xr, -q * yi);                                                            \
        x.y = __builtin_fma(p, xi, q * yr);                                                             \
        y.x = __builtin_fma(p, yr, -q * xi);                                                            \
        y.y = __builtin_fma(p, yi, q * xr);                                                             \
    }                                                                                                   \
    __device__ __forceinline__ void r_ry(v2d& x, v2d& y, double p, double q) {                          \
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;                                            \
        x.x = __builtin_fma(p, xr, -q * yr);                                                            \
        x.y = __builtin_fma(p, xi, -q * yi);                                                            \
        y.x = __builtin_fma(p, yr, q * xr);                                                             \
        y.y = __builtin_fma(p, yi, q * xi);                                                             \
    }                                                                                                   \
    __device__ __forceinline__ void r_sc(v2d& v, double dr, double di) {                                \
        const double pr = -di * v.y, pi = di * v.x;                                                     \
        v.x = __builtin_fma(dr, v.x, pr);                                                               \
        v.y = __builtin_fma(dr, v.y, pi);                                                               \
    }                                                                                                   \
    __device__ __forceinline__ v2d r_ng(v2d v, unsigned int sgn) {                                      \
        const unsigned long long s = (unsigned long long)sgn << 32;                                     \
        const double re = v.x, im = v.y;                                                                \
        v2d r;                                                                                          \
        r.x = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, re) ^ s);               \
        r.y = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, im) ^ s);               \
        return r;                                                                                       \
    }                                                                                                   \
    __device__ __forceinline__ void r_phase(double& sr, double& si, double dr, double di) {             \
        const double nr = __builtin_fma(sr, dr, -si * di);                                              \
        si = __builtin_fma(sr, di, si * dr);                                                            \
        sr = nr;                                                                                        \
    }                                                                                                   \
    __device__ __forceinline__ void r_shy(v2d& x, v2d& y, double a, double b) {                         \
        x.x = __builtin_fma(a, y.x, x.x);                                                               \
        x.y = __builtin_fma(a, y.y, x.y);                                                               \
        y.x = __builtin_fma(b, x.x, y.x);                                                               \
        y.y = __builtin_fma(b, x.y, y.y);                                                               \
        x.x = __builtin_fma(a, y.x, x.x);                                                               \
        x.y = __builtin_fma(a, y.y, x.y);                                                               \
    }                                                                                                   \
    __device__ __forceinline__ void r_shx(v2d& x, v2d& y, double a, double b) {                         \
        x.x = __builtin_fma(-a, y.y, x.x);                                                              \
        x.y = __builtin_fma(a, y.x, x.y);                                                               \
        y.x = __builtin_fma(-b, x.y, y.x);                                                              \
        y.y = __builtin_fma(b, x.x, y.y);                                                               \
        x.x = __builtin_fma(-a, y.y, x.x);                                                              \
        x.y = __builtin_fma(a, y.x, x.y);                                                               \
    }                                                                                                   \
    __device__ __forceinline__ void r_nu0(v2d& x, v2d& y, double br, double bi, double cr, double ci) { \
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;                                            \
        x.x = __builtin_fma(-bi, yi, __builtin_fma(br, yr, xr));                                        \
        x.y = __builtin_fma(bi, yr, __builtin_fma(br, yi, xi));                                         \
        y.x = __builtin_fma(-ci, xi, __builtin_fma(cr, xr, yr));                                        \
        y.y = __builtin_fma(ci, xr, __builtin_fma(cr, xi, yi));                                         \
    }                                                                                                   \
    __device__ __forceinline__ void r_nu1(v2d& x, v2d& y, double br, double bi, double cr, double ci) { \
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;                                            \
        x.x = __builtin_fma(-bi, xi, __builtin_fma(br, xr, yr));                                        \
        x.y = __builtin_fma(bi, xr, __builtin_fma(br, xi, yi));                                         \
        y.x = __builtin_fma(-ci, yi, __builtin_fma(cr, yr, xr));                                        \
        y.y = __builtin_fma(ci, yr, __builtin_fma(cr, yi, xi));                                         \
    }

#define LLSV_STRINGIZE2(...) #__VA_ARGS__
#define LLSV_STRINGIZE(x) LLSV_STRINGIZE2(x)
