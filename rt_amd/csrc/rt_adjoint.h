// rt_adjoint.h -- reverse-mode derivatives of one surface of the trace.
//
// surface_adjoint() takes a ray's state on entry to a surface (before the
// frame transform) and the adjoint of its state after the surface, recomputes
// the surface's stages (transform, normalize, intersection, action,
// normalize; Scheme::run order, common-lib/scheme.cpp:983-1046) in plain FP64
// from the entry state, and runs them backwards: it returns the adjoint of the
// entry state and adds the derivatives with respect to the surface's
// parameters (curvature, conic constant, aspheric coefficients, toric cR,
// decentres, tilts, grating constant) to gp[].  The loss is a function of the
// surviving rays' positions; vignetting, the choice of root and the Newton
// iteration count are held fixed, as they are piecewise constant.
//
// Intersections: the plane and the closed-form conic (quadric) roots through
// the implicit function F(p + s c) = 0; the iterative conic and toric
// intersections (rt_cuda_engine_surface.cu:257-333, 414-492) return the
// penultimate Newton iterate, so their derivative is taken through the
// iterations themselves, s_{k+1} = s_k - F(s_k)/F'(s_k), with the iteration
// count of the forward pass.
//
// Everything here is plain arithmetic on doubles (host and device), so it
// also builds for the host, where it is checked against finite differences.
#pragma once

#include "rt_hip.h"

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#include "rt_device.h"     // the refined reciprocal / reciprocal square root of the device build
#define RT_HD __host__ __device__ inline
#define RT_HDM __host__ __device__
#define RT_UNROLL _Pragma("unroll")
#else
#include <math.h>
#define RT_HD static inline
#define RT_HDM
#define RT_UNROLL
#endif

namespace rt {
namespace adj {

// The derivative formulas below run only in the reverse pass (the forward
// pass is the trace's own stage code, compiled with multiply-adds fused
// within a source expression only): here the compiler may fuse a multiply
// and an add across statements too.
#if defined(__clang__)
#pragma clang fp contract(fast)
#endif

// reciprocal and reciprocal square root: the refined hardware estimates of
// rt_device.h on the device, IEEE on the host
RT_HD double drcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rt::rcp(x);
#else
    return 1.0 / x;
#endif
}

RT_HD double drsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rt::rsqrt(x);
#else
    return 1.0 / sqrt(x);
#endif
}

RT_HD double dsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rt::fsqrt(x);
#else
    return sqrt(x);
#endif
}

// 1/sqrt(s) for the squared norm s of a direction: IEEE on the host; on the
// device the form of the trace's normalization (rt::normalize_ray): for
// s = 1 + d with |d| < 1e-8 on every lane of the wave 1 - d/2 (within 4e-17
// of 1/sqrt(s)), else the refined reciprocal square root
RT_HD double dinvnorm(double s) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double d = s - 1.0;
    const bool near = fabs(d) < 1e-8;
    if (__all(near)) return fma(-0.5, d, 1.0);
    return near ? fma(-0.5, d, 1.0) : rt::rsqrt(s);
#else
    return 1.0 / sqrt(s);
#endif
}

// parameter slots of one surface
enum ParamSlot {
    PG_C = 0,                          // curvature c
    PG_K = 1,                          // conic constant k
    PG_A = 2,                          // aspheric a_1 .. a_RT_MAX_POLY (coefficients of r^2 .. r^2N)
    PG_CR = PG_A + RT_MAX_POLY,        // toric cR
    PG_DIST = PG_CR + 1,               // decentres d_x, d_y, d_z
    PG_ANG = PG_DIST + 3,              // tilts about Y, X, Z (radians, as in the record)
    PG_D = PG_ANG + 3,                 // grating constant (um)
    PG_N = PG_D + 1
};

// public record field -> slot (-1: not differentiable)
RT_HD int slot_of_field(int field) {
    if (field == RT_S_C) return PG_C;
    if (field == RT_S_K) return PG_K;
    if (field >= RT_S_A && field < RT_S_A + RT_MAX_POLY) return PG_A + (field - RT_S_A);
    if (field == RT_S_CR) return PG_CR;
    if (field >= RT_S_DIST && field < RT_S_DIST + 3) return PG_DIST + (field - RT_S_DIST);
    if (field >= RT_S_ANG && field < RT_S_ANG + 3) return PG_ANG + (field - RT_S_ANG);
    if (field == RT_S_GRCONST) return PG_D;
    return -1;
}

struct St {
    double X, Y, Z, cX, cY, cZ;
};

enum { AK_PLANE = 0, AK_QUAD = 1, AK_CONIC_NEWTON = 2, AK_TORIC_NEWTON = 3, AK_NONE = 4 };

// The structure of a surface (class, type, number of aspheric terms, grating
// type, intersection kind, which rotations are applied) comes from a record
// view: DynRec reads it from the record at run time (host, generic kernel);
// the generated kernels pass rt::StatSurf<...> (rt_device.h), whose structure
// is compile-time, so every branch below on it folds away and the parameter
// slots are fixed indices.  Values are always read through .p.
struct DynRec {
    const double* p;
    RT_HDM int cls() const { return (int)p[RT_S_CLASS]; }
    RT_HDM int type() const { return (int)p[RT_S_TYPE]; }
    RT_HDM int npoly() const { return (int)p[RT_S_NPOLY]; }
    RT_HDM int grtype() const { return (int)p[RT_S_GRTYPE]; }
    RT_HDM bool rot(int j) const { return p[RT_S_ANG + j] != 0.0; }
    RT_HDM bool trans() const {
        return (p[RT_S_DIST + 2] != 0.0) || (p[RT_S_DIST + 0] != 0.0) || (p[RT_S_DIST + 1] != 0.0);
    }
    RT_HDM int ikind() const {
        const int c = cls(), n = npoly();
        if (c == RT_PLANE || (c == RT_CONIC && n == 0 && p[RT_S_C] == 0.0)) return AK_PLANE;
        if (c == RT_CONIC) return n == 0 ? AK_QUAD : AK_CONIC_NEWTON;
        if (c == RT_TORIC) return AK_TORIC_NEWTON;
        return AK_NONE;
    }
};

// ------------------------------------------------------------ polynomials
// P(x) = sum a_i x^i, Q = P'(x) = sum i a_i x^(i-1), R = Q'(x)
struct Poly {
    double P, Q, R;
};

template <class Rec>
RT_HD Poly poly(const Rec& d, double x) {
    const auto p = d.p;
    const int N = d.npoly();
    Poly o = {0.0, 0.0, 0.0};
    double xi = 1.0;                   // x^(i-1)
    double xm = 0.0;                   // x^(i-2)
    for (int i = 1; i <= N; ++i) {
        const double a = p[RT_S_A + i - 1];
        o.P += a * xi * x;
        o.Q += i * a * xi;
        if (i >= 2) o.R += i * (i - 1) * a * xm;
        xm = xi;
        xi *= x;
    }
    return o;
}

// d/d a_i of P and of Q at x, accumulated with weights wP, wQ
template <class Rec>
RT_HD void poly_param_adj(const Rec& d, double x, double wP, double wQ, double* gp) {
    const int N = d.npoly();
    double xi = 1.0;
    for (int i = 1; i <= N; ++i) {
        gp[PG_A + i - 1] += wP * xi * x + wQ * i * xi;
        xi *= x;
    }
}

// ---------------------------------------------------- conic profile pieces
// With t = 1 - (k+1) c^2 x, q = sqrt(t) (x = r^2, or Y^2 for a toric):
//   S(x) = c x / (1 + q)   sag of the conic, dS/dx = c / (2q)
//   E(x) = c / q + 2 Q(x)  the factor of the conic normal, -dF/dX = E X
struct Conic {
    double q, S, E, dSdx, dEdx, dSdc, dSdk, dEdc, dEdk;
    Poly pl;
};

template <class Rec>
RT_HD Conic conic_at(const Rec& d, double x) {
    const auto p = d.p;
    const double c = p[RT_S_C], k1 = p[RT_S_K] + 1.0;
    Conic o;
    const double K = k1 * c * c;
    const double t = 1.0 - K * x;
    const double rq = drsqrt(t);               // 1/q
    o.q = t * rq;
    const double rq3 = rq * rq * rq, iopq = drcp(1.0 + o.q), iopq2 = iopq * iopq;
    o.pl = poly(d, x);
    o.S = c * x * iopq;
    o.E = c * rq + 2.0 * o.pl.Q;
    o.dSdx = 0.5 * c * rq;
    o.dEdx = 0.5 * c * K * rq3 + 2.0 * o.pl.R;
    o.dSdc = x * iopq + k1 * c * c * x * x * rq * iopq2;
    o.dSdk = 0.5 * c * c * c * x * x * rq * iopq2;
    o.dEdc = rq + k1 * c * c * x * rq3;
    o.dEdk = 0.5 * c * c * c * x * rq3;
    return o;
}

// adjoint of (S + P) and of E at x, with weights wS (on S + P: the full sag) and wE
template <class Rec>
RT_HD double conic_adj(const Rec& d, const Conic& k, double x, double wS, double wE, double* gp) {
    gp[PG_C] += wS * k.dSdc + wE * k.dEdc;
    gp[PG_K] += wS * k.dSdk + wE * k.dEdk;
    poly_param_adj(d, x, wS, 2.0 * wE, gp);
    return wS * (k.dSdx + k.pl.Q) + wE * k.dEdx;      // adjoint of x
}

// --------------------------------------------------------------- transform
RT_HD void rot2(double s, double c, double& v1, double& v2) {
    double e = v1 * c - v2 * s;
    v2 = v1 * s + v2 * c;
    v1 = e;
}

template <class Rec>
RT_HD void fwd_transform(const Rec& d, St& r) {
    const auto p = d.p;
    if (d.trans()) {
        r.X -= p[RT_S_DIST + 0]; r.Y -= p[RT_S_DIST + 1]; r.Z -= p[RT_S_DIST + 2];
    }
    if (d.rot(0)) { rot2(p[RT_S_SIN + 0], p[RT_S_COS + 0], r.X, r.Z); rot2(p[RT_S_SIN + 0], p[RT_S_COS + 0], r.cX, r.cZ); }
    if (d.rot(1)) { rot2(p[RT_S_SIN + 1], p[RT_S_COS + 1], r.Y, r.Z); rot2(p[RT_S_SIN + 1], p[RT_S_COS + 1], r.cY, r.cZ); }
    if (d.rot(2)) { rot2(p[RT_S_SIN + 2], p[RT_S_COS + 2], r.X, r.Y); rot2(p[RT_S_SIN + 2], p[RT_S_COS + 2], r.cX, r.cY); }
}

// one rotation backwards: outputs (v1, v2) become inputs, adjoints (w1, w2)
// likewise; the derivative with respect to the angle is added to *ga.  A zero
// angle (skipped forward) still has its derivative, the rotation's at 0.
RT_HD void rot2_adj(double s, double c, double& v1, double& v2, double& w1, double& w2, double* ga) {
    *ga += -w1 * v2 + w2 * v1;
    double i1 = c * v1 + s * v2, i2 = -s * v1 + c * v2;
    double a1 = c * w1 + s * w2, a2 = -s * w1 + c * w2;
    v1 = i1; v2 = i2; w1 = a1; w2 = a2;
}

// r: state after the transform (becomes the state before it); a: its adjoint
template <class Rec>
RT_HD void adj_transform(const Rec& d, St& r, St& a, double* gp) {
    const auto p = d.p;
    const double s2 = d.rot(2) ? p[RT_S_SIN + 2] : 0.0, c2 = d.rot(2) ? p[RT_S_COS + 2] : 1.0;
    const double s1 = d.rot(1) ? p[RT_S_SIN + 1] : 0.0, c1 = d.rot(1) ? p[RT_S_COS + 1] : 1.0;
    const double s0 = d.rot(0) ? p[RT_S_SIN + 0] : 0.0, c0 = d.rot(0) ? p[RT_S_COS + 0] : 1.0;
    double ga2 = 0.0, ga1 = 0.0, ga0 = 0.0;
    rot2_adj(s2, c2, r.X, r.Y, a.X, a.Y, &ga2);
    rot2_adj(s2, c2, r.cX, r.cY, a.cX, a.cY, &ga2);
    rot2_adj(s1, c1, r.Y, r.Z, a.Y, a.Z, &ga1);
    rot2_adj(s1, c1, r.cY, r.cZ, a.cY, a.cZ, &ga1);
    rot2_adj(s0, c0, r.X, r.Z, a.X, a.Z, &ga0);
    rot2_adj(s0, c0, r.cX, r.cZ, a.cX, a.cZ, &ga0);
    gp[PG_ANG + 0] += ga0;
    gp[PG_ANG + 1] += ga1;
    gp[PG_ANG + 2] += ga2;
    gp[PG_DIST + 0] -= a.X;
    gp[PG_DIST + 1] -= a.Y;
    gp[PG_DIST + 2] -= a.Z;
    r.X += p[RT_S_DIST + 0]; r.Y += p[RT_S_DIST + 1]; r.Z += p[RT_S_DIST + 2];
}

// --------------------------------------------------------------- normalize
RT_HD void fwd_normalize(St& r) {
    const double in = dinvnorm(r.cX * r.cX + r.cY * r.cY + r.cZ * r.cZ);
    r.cX *= in; r.cY *= in; r.cZ *= in;
}

// c: direction before normalizing; a: adjoint of the normalized direction (in/out)
RT_HD void adj_normalize(double cX, double cY, double cZ, double& aX, double& aY, double& aZ) {
    const double in = dinvnorm(cX * cX + cY * cY + cZ * cZ);
    const double uX = cX * in, uY = cY * in, uZ = cZ * in;
    const double d = uX * aX + uY * aY + uZ * aZ;
    aX = (aX - uX * d) * in;
    aY = (aY - uY * d) * in;
    aZ = (aZ - uZ * d) * in;
}

// ------------------------------------------------------------ intersection
constexpr int kMaxNewtonAdj = 8;    // iterations differentiated exactly (the loops below are unrolled
                                    // over this bound, so the iterates stay in registers)

struct NewtonSeq {
    int n;                          // iterations of the forward pass
    double s0;
    double s[kMaxNewtonAdj];        // s^(0) .. s^(n-1); the point is at s^(n-1)
};

// The point at parameter s along the ray, the profile there, and F and
// dF/ds of the iterative intersections
struct NewtonAt {
    double X, Y, Z, F, G;
    Conic k;
};

template <class Rec>
RT_HD NewtonAt newton_at(const Rec& d, int kind, const St& in, double s0, double s) {
    const auto p = d.p;
    NewtonAt o;
    const double u = s + s0;
    o.X = in.X + u * in.cX; o.Y = in.Y + u * in.cY; o.Z = s * in.cZ;
    if (kind == AK_CONIC_NEWTON) {
        o.k = conic_at(d, o.X * o.X + o.Y * o.Y);
        o.F = o.Z - o.k.S - o.k.pl.P;
        o.G = -o.k.E * o.X * in.cX - o.k.E * o.Y * in.cY + in.cZ;
    } else {
        const double cR = p[RT_S_CR];
        o.k = conic_at(d, o.Y * o.Y);
        const double Fy = o.k.S + o.k.pl.P;
        const double dFy = o.k.E * o.Y;
        o.F = o.Z - Fy - 0.5 * cR * (o.X * o.X + o.Z * o.Z - Fy * Fy);
        o.G = -cR * o.X * in.cX + (cR * Fy - 1.0) * dFy * in.cY + (1.0 - cR * o.Z) * in.cZ;
    }
    return o;
}

// F and dF/ds of the iterative intersections at parameter s along the ray
template <class Rec>
RT_HD void newton_FG(const Rec& d, int kind, const St& in, double s0, double s, double& F, double& G) {
    const NewtonAt o = newton_at(d, kind, in, s0, s);
    F = o.F;
    G = o.G;
}

// forward intersection from `in`; n_fwd > 0 replays that many Newton
// iterations, 0 iterates to the reference's convergence test; `given` (the
// forward pass's own iterates s^(0) .. s^(n_fwd - 1), as its tape holds them)
// takes their place without replaying them.  false: lost, or more iterations
// than are differentiated here
template <class Rec>
RT_HD bool fwd_intersect(const Rec& d, const St& in, int n_fwd, St& out, NewtonSeq& q,
                         const NewtonSeq* given = nullptr) {
    const auto p = d.p;
    const int kind = d.ikind();
    out = in;
    q.n = 0;
    q.s0 = 0.0;
    if (kind == AK_PLANE) {
        const double s0 = -in.Z * drcp(in.cZ);
        if (s0 < 0.0) return false;
        out.X = in.X + in.cX * s0; out.Y = in.Y + in.cY * s0; out.Z = 0.0;
        return true;
    }
    if (kind == AK_QUAD) {
        const double c = p[RT_S_C], k1 = p[RT_S_K] + 1.0, ic = drcp(c);
        const double A = k1 * in.cZ * in.cZ + in.cX * in.cX + in.cY * in.cY;
        const double B = 2.0 * (k1 * in.Z * in.cZ + in.X * in.cX + in.Y * in.cY - in.cZ * ic);
        const double C = k1 * in.Z * in.Z + in.X * in.X + in.Y * in.Y - 2.0 * in.Z * ic;
        const double D = B * B - 4.0 * A * C;
        if (D < 0.0) return false;
        const double sq = dsqrt(D), sgnB = (B < 0.0) ? -1.0 : 1.0, sigma = (c > 0) ? -1.0 : 1.0;
        const double qq = -0.5 * (B + sgnB * sq);
        double s;
        if (sigma == sgnB) s = (qq != 0.0) ? C * drcp(qq) : (-B + sigma * sq) * drcp(2.0 * A);
        else s = qq * drcp(A);
        out.X = in.X + s * in.cX; out.Y = in.Y + s * in.cY; out.Z = in.Z + s * in.cZ;
        return true;
    }
    if ((kind == AK_CONIC_NEWTON || kind == AK_TORIC_NEWTON) && given) {
        q.s0 = -in.Z * drcp(in.cZ);
        if (!(q.s0 >= 0) || n_fwd < 1 || n_fwd > kMaxNewtonAdj) return false;
        double s1 = 0.0;
RT_UNROLL
        for (int k = 0; k < kMaxNewtonAdj; ++k) {
            q.s[k] = given->s[k];
            if (k == n_fwd - 1) s1 = q.s[k];
        }
        q.n = n_fwd;
        const double u = s1 + q.s0;
        out.X = in.X + u * in.cX; out.Y = in.Y + u * in.cY; out.Z = s1 * in.cZ;
        return true;
    }
    if (kind == AK_CONIC_NEWTON || kind == AK_TORIC_NEWTON) {
        q.s0 = -in.Z * drcp(in.cZ);
        if (!(q.s0 >= 0)) return false;
        double s1 = 0.0, s2 = 0.0;
        int it = 0;
        bool done = false;
RT_UNROLL
        for (int k = 0; k < kMaxNewtonAdj; ++k) {
            if (!done) {
                s1 = s2;
                q.s[k] = s1;
                double F, G;
                newton_FG(d, kind, in, q.s0, s1, F, G);
                if (fabs(G) < 1.0e-7) return false;
                s2 = s1 - F * drcp(G);
                it = k + 1;
                done = n_fwd > 0 ? it >= n_fwd : !(fabs(s2 - s1) >= 1.0e-7);
            }
        }
        if (!done) return false;            // more iterations than are differentiated here
        q.n = it;
        const double u = s1 + q.s0;
        out.X = in.X + u * in.cX; out.Y = in.Y + u * in.cY; out.Z = s1 * in.cZ;
        return true;
    }
    return true;
}

// adjoint of the point at parameter s (X = X0 + (s + s0) cX, Y = .., Z = s cZ)
RT_HD void point_adj(const St& in, double s0, double s, double aX, double aY, double aZ, St& ain, double& as,
                     double& as0) {
    const double u = s + s0;
    const double au = aX * in.cX + aY * in.cY;
    as += au + aZ * in.cZ;
    as0 += au;
    ain.X += aX;
    ain.Y += aY;
    ain.cX += aX * u;
    ain.cY += aY * u;
    ain.cZ += aZ * s;
}

// adjoint of F, G (newton_at, evaluated as `o`): point adjoint out
// (aX, aY, aZ), direct direction adjoints into ain, parameters into gp
template <class Rec>
RT_HD void newton_FG_adj(const Rec& d, int kind, const St& in, const NewtonAt& o, double aF, double aG,
                         double& aX, double& aY, double& aZ, St& ain, double* gp) {
    const auto p = d.p;
    const double X = o.X, Y = o.Y, Z = o.Z;
    const Conic& k = o.k;
    aX = aY = aZ = 0.0;
    if (kind == AK_CONIC_NEWTON) {
        const double r2 = X * X + Y * Y;
        const double W = X * in.cX + Y * in.cY;
        // G = -E W + cZ
        const double aE = -aG * W, aW = -aG * k.E;
        ain.cZ += aG;
        aX += aW * in.cX; ain.cX += aW * X;
        aY += aW * in.cY; ain.cY += aW * Y;
        // F = Z - (S + P)
        aZ += aF;
        const double ar2 = conic_adj(d, k, r2, -aF, aE, gp);
        aX += 2.0 * X * ar2;
        aY += 2.0 * Y * ar2;
    } else {
        const double cR = p[RT_S_CR];
        const double Y2 = Y * Y;
        const double Fy = k.S + k.pl.P, dFy = k.E * Y;
        // G = -cR X cX + (cR Fy - 1) dFy cY + (1 - cR Z) cZ
        aX += -aG * cR * in.cX; ain.cX += -aG * cR * X;
        double aFy = aG * cR * dFy * in.cY;
        double adFy = aG * (cR * Fy - 1.0) * in.cY;
        ain.cY += aG * (cR * Fy - 1.0) * dFy;
        aZ += -aG * cR * in.cZ; ain.cZ += aG * (1.0 - cR * Z);
        gp[PG_CR] += aG * (-X * in.cX + Fy * dFy * in.cY - Z * in.cZ);
        // F = Z - Fy - cR/2 (X^2 + Z^2 - Fy^2)
        aZ += aF * (1.0 - cR * Z);
        aX += aF * (-cR * X);
        aFy += aF * (-1.0 + cR * Fy);
        gp[PG_CR] += aF * (-0.5 * (X * X + Z * Z - Fy * Fy));
        // Fy(Y) = S + P at Y^2; dFy(Y) = E Y
        aY += aFy * dFy + adFy * (k.E + 2.0 * Y2 * k.dEdx);
        const double aY2 = conic_adj(d, k, Y2, aFy, adFy * Y, gp);
        (void)aY2;   // the Y^2 chain of Fy and E is already in dFy and dE/dY above
    }
}

// in: state before the intersection; out: the intersection point; a: adjoint of
// `out` (becomes the adjoint of `in`)
template <class Rec>
RT_HD void adj_intersect(const Rec& d, const St& in, const St& out, const NewtonSeq& q, St& a, double* gp) {
    const auto p = d.p;
    const int kind = d.ikind();
    St ain = {0, 0, 0, a.cX, a.cY, a.cZ};
    if (kind == AK_PLANE) {
        const double icZ = drcp(in.cZ);
        const double s0 = -in.Z * icZ;
        const double as0 = a.X * in.cX + a.Y * in.cY;
        ain.X = a.X; ain.Y = a.Y;
        ain.cX += a.X * s0; ain.cY += a.Y * s0;
        ain.Z = -as0 * icZ;
        ain.cZ += as0 * in.Z * icZ * icZ;
    } else if (kind == AK_QUAD) {
        const double c = p[RT_S_C], k1 = p[RT_S_K] + 1.0, ic = drcp(c);
        const double s = ((out.X - in.X) * in.cX + (out.Y - in.Y) * in.cY + (out.Z - in.Z) * in.cZ) *
                         drcp(in.cX * in.cX + in.cY * in.cY + in.cZ * in.cZ);
        const double gX = 2.0 * out.X, gY = 2.0 * out.Y, gZ = 2.0 * k1 * out.Z - 2.0 * ic;
        const double gc = gX * in.cX + gY * in.cY + gZ * in.cZ;
        const double f = (a.X * in.cX + a.Y * in.cY + a.Z * in.cZ) * drcp(gc);
        ain.X = a.X - f * gX; ain.Y = a.Y - f * gY; ain.Z = a.Z - f * gZ;
        ain.cX += s * (a.X - f * gX); ain.cY += s * (a.Y - f * gY); ain.cZ += s * (a.Z - f * gZ);
        gp[PG_K] += -f * out.Z * out.Z;
        gp[PG_C] += -f * 2.0 * out.Z * ic * ic;
    } else if (kind == AK_CONIC_NEWTON || kind == AK_TORIC_NEWTON) {
        double as = 0.0, as0 = 0.0;
        const int n = q.n;
RT_UNROLL
        for (int k = kMaxNewtonAdj - 1; k >= 0; --k) {
            if (k == n - 1) {
                // the returned point, at the last iterate evaluated
                point_adj(in, q.s0, q.s[k], a.X, a.Y, a.Z, ain, as, as0);
            } else if (k < n - 1) {
                // s^(k+1) = s^(k) - F/G at s^(k)
                const NewtonAt o = newton_at(d, kind, in, q.s0, q.s[k]);
                const double iG = drcp(o.G);
                const double aF = -as * iG, aG = as * o.F * iG * iG;
                double pX, pY, pZ;
                newton_FG_adj(d, kind, in, o, aF, aG, pX, pY, pZ, ain, gp);
                point_adj(in, q.s0, q.s[k], pX, pY, pZ, ain, as, as0);   // as keeps the direct term
            }
        }
        // s0 = -Z / cZ
        const double icZ = drcp(in.cZ);
        ain.Z += -as0 * icZ;
        ain.cZ += as0 * in.Z * icZ * icZ;
    } else {
        ain.X = a.X; ain.Y = a.Y; ain.Z = a.Z;
    }
    a = ain;
}

// ------------------------------------------------------------------ normals
// Surface normal (unnormalised, the engine's convention) at a point and the
// adjoint of a normal back to the point and the parameters.
struct Normal {
    double x, y, z;
};

// The normal at a point with the profile it was formed from (evaluated once
// for the action, its adjoint and the normal's adjoint)
struct Geom {
    Normal N;
    Conic k;
};

template <class Rec>
RT_HD Geom geom_at(const Rec& d, double X, double Y, double Z) {
    const auto p = d.p;
    const int cls = d.cls();
    Geom g;
    if (cls == RT_CONIC) {
        g.k = conic_at(d, X * X + Y * Y);
        g.N = Normal{-g.k.E * X, -g.k.E * Y, 1.0};
    } else if (cls == RT_TORIC) {
        const double cR = p[RT_S_CR];
        g.k = conic_at(d, Y * Y);
        const double Fy = g.k.S + g.k.pl.P, dFy = g.k.E * Y;
        g.N = Normal{-cR * X, (cR * Fy - 1.0) * dFy, 1.0 - cR * Z};
    } else {
        g.N = Normal{0.0, 0.0, 1.0};
    }
    return g;
}

template <class Rec>
RT_HD void normal_adj(const Rec& d, double X, double Y, double Z, const Conic& k, const Normal& aN, St& apt,
                      double* gp) {
    const auto p = d.p;
    const int cls = d.cls();
    if (cls == RT_CONIC) {
        const double r2 = X * X + Y * Y;
        const double aE = -(aN.x * X + aN.y * Y);
        apt.X += -k.E * aN.x;
        apt.Y += -k.E * aN.y;
        const double ar2 = conic_adj(d, k, r2, 0.0, aE, gp);
        apt.X += 2.0 * X * ar2;
        apt.Y += 2.0 * Y * ar2;
    } else if (cls == RT_TORIC) {
        const double cR = p[RT_S_CR];
        const double Y2 = Y * Y;
        const double Fy = k.S + k.pl.P, dFy = k.E * Y;
        apt.X += -cR * aN.x;
        apt.Z += -cR * aN.z;
        gp[PG_CR] += -X * aN.x + Fy * dFy * aN.y - Z * aN.z;
        const double aFy = aN.y * cR * dFy, adFy = aN.y * (cR * Fy - 1.0);
        apt.Y += aFy * dFy + adFy * (k.E + 2.0 * Y2 * k.dEdx);
        (void)conic_adj(d, k, Y2, aFy, adFy * Y, gp);
    }
}

// ------------------------------------------------------------------ actions
// The action's output direction (before the normalisation that follows);
// false for total internal reflection or an evanescent order.
// (g: geom_at of the point, read for curved mirrors and lenses)
template <class Rec>
RT_HD bool fwd_action(const Rec& d, const St& r, const Geom& g, double mu, double L, double& oX, double& oY,
                      double& oZ) {
    const int type = d.type(), cls = d.cls();
    oX = r.cX; oY = r.cY; oZ = r.cZ;
    if (type == RT_MIRROR) {
        if (cls == RT_PLANE) { oZ = -r.cZ; return true; }
        const Normal N = g.N;
        const double a2 = 2.0 * (r.cX * N.x + r.cY * N.y + r.cZ * N.z) * drcp(N.x * N.x + N.y * N.y + N.z * N.z);
        oX = r.cX - a2 * N.x; oY = r.cY - a2 * N.y; oZ = r.cZ - a2 * N.z;
        return true;
    }
    if (type == RT_LENS) {
        if (cls == RT_PLANE) {
            const double a = mu * r.cZ, b = mu * mu - 1.0;
            if (b > a * a) return false;
            const double G = -a + dsqrt(a * a - b);
            oX = mu * r.cX; oY = mu * r.cY; oZ = mu * r.cZ + G;
            return true;
        }
        const Normal N = g.N;
        const double iM = drcp(N.x * N.x + N.y * N.y + N.z * N.z);
        const double a = mu * (r.cX * N.x + r.cY * N.y + r.cZ * N.z) * iM, b = (mu * mu - 1.0) * iM;
        if (b > a * a) return false;
        const double G = -a + dsqrt(a * a - b);
        oX = mu * r.cX + G * N.x; oY = mu * r.cY + G * N.y; oZ = mu * r.cZ + G * N.z;
        return true;
    }
    if (type == RT_GRATING) {
        if (cls != RT_PLANE) return false;                 // curved gratings: no adjoint
        if (d.grtype() == RT_REFLECTIVE) {
            const double a = r.cZ, b = L * L - 2.0 * L * r.cX;
            if (b > a * a) return false;
            const double G = -a - dsqrt(a * a - b);
            oX = r.cX - L; oZ = r.cZ + G;
        } else {
            const double a = mu * r.cZ, b = mu * mu - 1.0 + L * L - 2.0 * mu * L * r.cX;
            if (b > a * a) return false;
            const double G = -a + dsqrt(a * a - b);
            oX = mu * r.cX - L; oY = mu * r.cY; oZ = mu * r.cZ + G;
        }
        return true;
    }
    return true;
}

// r: state on the surface before the action; (aX, aY, aZ): adjoint of the
// action's output direction; a: receives the adjoint of r (its point part is
// added to, its direction part set); aL: adjoint of L
template <class Rec>
RT_HD bool adj_action(const Rec& d, const St& r, const Geom& g, double mu, double L, double aX, double aY,
                      double aZ, St& a, double& aL, double* gp) {
    const int type = d.type(), cls = d.cls();
    aL = 0.0;
    if (type == RT_MIRROR) {
        if (cls == RT_PLANE) { a.cX = aX; a.cY = aY; a.cZ = -aZ; return true; }
        const Normal N = g.N;
        const double D = r.cX * N.x + r.cY * N.y + r.cZ * N.z, M = N.x * N.x + N.y * N.y + N.z * N.z;
        const double aoN = aX * N.x + aY * N.y + aZ * N.z;
        const double iM = drcp(M), q2 = -2.0 * D * iM;
        const double aD = -2.0 * iM * aoN, aM = 2.0 * D * iM * iM * aoN;
        Normal aN = {q2 * aX + aD * r.cX + 2.0 * aM * N.x, q2 * aY + aD * r.cY + 2.0 * aM * N.y,
                     q2 * aZ + aD * r.cZ + 2.0 * aM * N.z};
        a.cX = aX + aD * N.x; a.cY = aY + aD * N.y; a.cZ = aZ + aD * N.z;
        normal_adj(d, r.X, r.Y, r.Z, g.k, aN, a, gp);
        return true;
    }
    if (type == RT_LENS) {
        if (cls == RT_PLANE) {
            const double av = mu * r.cZ, b = mu * mu - 1.0, sq = dsqrt(av * av - b);
            const double aa = aZ * (-1.0 + av * drcp(sq));
            a.cX = mu * aX; a.cY = mu * aY; a.cZ = mu * aZ + aa * mu;
            return true;
        }
        const Normal N = g.N;
        const double M = N.x * N.x + N.y * N.y + N.z * N.z;
        const double D = r.cX * N.x + r.cY * N.y + r.cZ * N.z;
        const double iM = drcp(M);
        const double av = mu * D * iM, b = (mu * mu - 1.0) * iM, sq = dsqrt(av * av - b), G = -av + sq;
        const double isq = drcp(sq);
        const double aG = aX * N.x + aY * N.y + aZ * N.z;
        const double aa = aG * (-1.0 + av * isq), ab = -0.5 * aG * isq;
        const double aD = aa * mu * iM, aM = (-aa * av - ab * b) * iM;
        Normal aN = {G * aX + aD * r.cX + 2.0 * aM * N.x, G * aY + aD * r.cY + 2.0 * aM * N.y,
                     G * aZ + aD * r.cZ + 2.0 * aM * N.z};
        a.cX = mu * aX + aD * N.x; a.cY = mu * aY + aD * N.y; a.cZ = mu * aZ + aD * N.z;
        normal_adj(d, r.X, r.Y, r.Z, g.k, aN, a, gp);
        return true;
    }
    if (type == RT_GRATING) {
        if (cls != RT_PLANE) return false;
        if (d.grtype() == RT_REFLECTIVE) {
            const double av = r.cZ, b = L * L - 2.0 * L * r.cX, isq = drcp(dsqrt(av * av - b));
            const double aG = aZ, aa = aG * (-1.0 - av * isq), ab = 0.5 * aG * isq;
            a.cX = aX + ab * (-2.0 * L);
            a.cY = aY;
            a.cZ = aZ + aa;
            aL = -aX + ab * (2.0 * L - 2.0 * r.cX);
        } else {
            const double av = mu * r.cZ, b = mu * mu - 1.0 + L * L - 2.0 * mu * L * r.cX;
            const double isq = drcp(dsqrt(av * av - b));
            const double aG = aZ, aa = aG * (-1.0 + av * isq), ab = -0.5 * aG * isq;
            a.cX = mu * aX + ab * (-2.0 * mu * L);
            a.cY = mu * aY;
            a.cZ = mu * aZ + aa * mu;
            aL = -aX + ab * (2.0 * L - 2.0 * mu * r.cX);
        }
        return true;
    }
    a.cX = aX; a.cY = aY; a.cZ = aZ;
    return true;
}

// ---------------------------------------------------------------- surface
// entry: ray state on entry to the surface (previous frame); n_fwd: Newton
// iterations of the forward pass (0: recompute to convergence); mu = n1/n2
// and L = m lambda / n2 / d of the forward pass (lenses, gratings); a: adjoint
// of the state after the surface (in), of the entry state (out).  false when
// the surface has no adjoint here (curved gratings) or the replayed ray is lost.
template <class Rec>
RT_HD bool surface_adjoint(const Rec& d, const St& entry, int n_fwd, double mu, double L, St& a, double* gp,
                           const NewtonSeq* given = nullptr) {
    const auto p = d.p;
    const int type = d.type();
    St sT = entry;
    fwd_transform(d, sT);
    St sN = sT;
    fwd_normalize(sN);
    if (type == RT_AUXTYPE) {
        adj_normalize(sT.cX, sT.cY, sT.cZ, a.cX, a.cY, a.cZ);
        adj_transform(d, sT, a, gp);
        return true;
    }
    St sI;
    NewtonSeq q;
    if (!fwd_intersect(d, sN, n_fwd, sI, q, given)) return false;
    // the profile at the intersection, for curved mirrors and lenses
    Geom g{};
    if ((type == RT_MIRROR || type == RT_LENS) && d.cls() != RT_PLANE) g = geom_at(d, sI.X, sI.Y, sI.Z);
    double oX, oY, oZ;
    if (!fwd_action(d, sI, g, mu, L, oX, oY, oZ)) return false;
    // normalisation after the action
    adj_normalize(oX, oY, oZ, a.cX, a.cY, a.cZ);
    // action: point adjoint accumulates, direction adjoint replaced
    St aI = {a.X, a.Y, a.Z, 0.0, 0.0, 0.0};
    double aL = 0.0;
    if (!adj_action(d, sI, g, mu, L, a.cX, a.cY, a.cZ, aI, aL, gp)) return false;
    if (type == RT_GRATING && L != 0.0) gp[PG_D] += aL * (-L * drcp(p[RT_S_GRCONST]));
    // intersection, first normalisation, transform
    adj_intersect(d, sN, sI, q, aI, gp);
    adj_normalize(sT.cX, sT.cY, sT.cZ, aI.cX, aI.cY, aI.cZ);
    adj_transform(d, sT, aI, gp);
    a = aI;
    return true;
}

#if defined(__clang__)
#pragma clang fp contract(on)
#endif

}  // namespace adj
}  // namespace rt

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
// ------------------------------------------------------------------ kernels
// The adjoint kernel body, shared by the generic kernel (rt_adjoint.hip,
// runtime surface descriptors) and the kernels generated per scheme structure
// and parameter list (rt_jit.cpp, compile-time descriptors and gradient rows).
// Each lane traces its rays forward through the scheme (the trace's own
// surface steps), keeping the state on entry to every surface and that
// surface's Newton iteration count in a tape slot of its own (n_surf x 7
// doubles per lane, reused ray after ray), then walks the surfaces backwards
// with surface_adjoint.  The loss enters as a seed per ray group
// g = ray mod n_groups: adjoint of the final (X, Y) = (ax_g + b_g X, ay_g + b_g Y).
// Parameter derivatives go to per-lane accumulators (n_par x lanes), reduced
// afterwards in a fixed order.
namespace rt {
namespace adj {

constexpr int kTape = 7;                 // X Y Z cX cY cZ, Newton iterations

struct AdjArgs {
    double beam[kBeamRec];               // beam fields + BeamUniform (first: kernarg offset 0)
    const double* table;                 // device records
    const double* pool;
    int n_surf, pool_len, n_par, pad;
    uint64_t seed, n, n_groups;
    const double* seeds;                 // n_groups x 3: ax, ay, b
    const int* prow;                     // n_surf x PG_N: the caller's index of (surface, slot), or -1
    double* tape;                        // (n_surf * kTape) x lanes
    double* gacc;                        // n_par x waves (single scheme); tile-wave x n_par (batch)
    unsigned long long* stats;
    // batched adjoint (adjoint_batch_body): instance i owns records
    // table[i * inst_stride ..], rays 0 .. ray_start[i+1] - ray_start[i] - 1 of
    // its own beam (seed + i), tiles [tile_start[i], tile_start[i+1]) of one
    // workgroup each, and group seeds seeds[(i * n_groups + g) * 3 ..]
    const uint64_t* tile_start;
    const uint64_t* ray_start;
    const uint32_t* tile_inst;           // instance of each tile
    uint64_t n_inst, n_tiles, inst_stride;
    // grid-wavelength beams: per-wavelength table of the spline-valued
    // functions (column q at ltab[q * lt_nk ..]); generated kernels built for it
    const double* ltab;
    uint64_t lt_nk;
    // batched adjoint with a beam per instance: n_inst records of
    // kBeamRec doubles (beam fields, then the uniform values), or null
    const double* inst_beams;
    // rows of the tape per lane: kTape per surface, and kMaxNewtonAdj more per
    // iterating intersection where the forward records its Newton iterates
    uint64_t tape_rows;
};

// Column q of the current ray's per-wavelength table values (AdjArgs is the
// adjoint kernels' only argument, at offset 0 of the argument segment, so
// the table pointer and size are read as scalars there).  The grid index is
// clamped into the table.
template <int NC>
__device__ __forceinline__ void load_lval_adj(double (&lv)[NC], const TraceCtx& c, int q) {
    typedef const __attribute__((address_space(4))) AdjArgs* cadj;
    cadj ka = (cadj)__builtin_amdgcn_kernarg_segment_ptr();
    __asm__ volatile("" : "+s"(ka));
    const uint64_t nk = ka->lt_nk;
    const uint64_t k = (uint64_t)c.lk < nk ? (uint64_t)c.lk : nk - 1;
    lv[q] = ka->ltab[(uint64_t)q * nk + k];
}

// Lane-interleaved rows of doubles (row k of lane l at (k * G + l) * 8 bytes)
// addressed as a buffer resource: the lane's byte offset stays in one VGPR
// and the row offset k * G * 8 is formed in an SGPR at the access, from a
// row size the caller re-reads per ray, so the compiler keeps no per-row
// address across the ray loop (held, the tape and accumulator addresses of
// a generated kernel would take more registers than a lane has).  Every
// access stays inside the resource's byte range, and a buffer access outside
// it would read 0 and write nothing.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kBufWord3 = 0x00020000;    // raw 32-bit dword access (gfx9 buffer resource word 3)

struct LaneRows {
    double* base;                        // wave-uniform
    int bytes;                           // byte range of the rows (wave-uniform)
    uint32_t voff;                       // lane * 8
    uint32_t row;                        // G * 8 (wave-uniform)
    // the resource and the row offset are wave-uniform; readfirstlane states
    // it where control flow with per-lane exits hides that from the compiler
    __device__ __amdgpu_buffer_rsrc_t rsrc() const {
        const uint64_t bp = (uint64_t)base;
        const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)bp);
        const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(bp >> 32));
        return __builtin_amdgcn_make_buffer_rsrc((double*)((hi << 32) | lo), (short)0,
                                                 __builtin_amdgcn_readfirstlane(bytes), kBufWord3);
    }
    __device__ uint32_t off(int k) const {
        return (uint32_t)__builtin_amdgcn_readfirstlane((int)((uint32_t)k * row));
    }
    __device__ double ld(int k) const {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc(), voff, off(k), 0));
    }
    __device__ void st(int k, double v) const {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rsrc(), voff, off(k), 0);
    }
    __device__ void add(int k, double v) const { st(k, ld(k) + v); }
};

__device__ __forceinline__ LaneRows lane_rows(double* base, uint64_t rows, uint64_t G, uint64_t lane) {
    LaneRows b;
    b.base = base;
    b.bytes = (int)(rows * G * sizeof(double));
    b.voff = (uint32_t)(lane * sizeof(double));
    uint32_t row = (uint32_t)(G * sizeof(double));
    row = (uint32_t)__builtin_amdgcn_readfirstlane((int)row);
    b.row = row;
    return b;
}

// forward through one surface: tape of the entry state (rows k0 .. k0+6),
// the trace's surface step, the Newton count
// (lv, col: the ray's per-wavelength table values, as surface_step takes them)
// (DERIV: lv holds the derived table's columns, as surface_step takes them;
// ITER: an iterating intersection records its first kMaxNewtonAdj Newton
// iterates in rows k0 + kTape ..)
struct TapeIterSink {
    const LaneRows* t;
    int k0;
    __device__ void operator()(unsigned k, double s) const {
        if (k < (unsigned)kMaxNewtonAdj) t->st(k0 + kTape + (int)k, s);
    }
};

template <bool DERIV = false, bool ITER = false, class SD>
__device__ __forceinline__ void fwd_step(const SD& d, int s, Ray& r, bool& alive, TraceCtx& c, const LaneRows& t,
                                         int k0, const double* lv = nullptr, int col = -1) {
    if (alive) {
        t.st(k0 + 0, r.X); t.st(k0 + 1, r.Y); t.st(k0 + 2, r.Z);
        t.st(k0 + 3, r.cX); t.st(k0 + 4, r.cY); t.st(k0 + 5, r.cZ);
    }
    if constexpr (ITER) surface_step<SD, 0, DERIV>(d, s, r, alive, c, lv, col, TapeIterSink{&t, k0});
    else surface_step<SD, 0, DERIV>(d, s, r, alive, c, lv, col);
    // the Newton count (row 6) only where the intersection iterates
    if ((d.ikind() == IK_CONIC_NEWTON || d.ikind() == IK_TORIC_NEWTON) && alive) t.st(k0 + 6, (double)c.last_it);
}

// backward through one surface: gp receives the surface's parameter
// derivatives (zero where the replay fails).  false: no adjoint (curved
// grating, replay lost).  The lanes of the wave are converged again when it
// returns, so the caller can reduce gp across the wave.  With lv (the ray's
// per-wavelength table values), n1 and n2 are lv[n1col] and lv[n2col], or
// the surface's constants where a column is -1.
// the Newton iterates the forward step recorded (fwd_step with ITER)
template <class SD>
__device__ __forceinline__ NewtonSeq taped_iterates(const SD&, const LaneRows& t, int k0, int nit) {
    NewtonSeq q;
    q.n = nit;
    q.s0 = 0.0;
RT_UNROLL
    for (int k = 0; k < kMaxNewtonAdj; ++k) q.s[k] = k < nit ? t.ld(k0 + kTape + k) : 0.0;
    return q;
}

template <bool ITER = false, class SD>
__device__ __forceinline__ bool bwd_step(const SD& d, const Ray& r, St& ad, const TraceCtx& c, const LaneRows& t,
                                         int k0, double* gp, const double* lv = nullptr, int n1col = -1,
                                         int n2col = -1) {
    const auto p = d.p;
RT_UNROLL
    for (int k = 0; k < PG_N; ++k) gp[k] = 0.0;
    const St entry = {t.ld(k0 + 0), t.ld(k0 + 1), t.ld(k0 + 2), t.ld(k0 + 3), t.ld(k0 + 4), t.ld(k0 + 5)};
    const int nit = (d.ikind() == IK_CONIC_NEWTON || d.ikind() == IK_TORIC_NEWTON) ? (int)t.ld(k0 + 6) : 0;
    // mu and L exactly as the forward pass formed them
    double mu = 0.0, L = 0.0;
    if (d.type() == RT_LENS || d.type() == RT_GRATING) {
        int e1 = 0, e2 = 0;
        double n1, n2;
        if (lv) {
            n1 = n1col >= 0 ? lv[n1col] : p[RT_S_N1VAL];
            n2 = n2col >= 0 ? lv[n2col] : p[RT_S_N2VAL];
        } else {
            n1 = fn_eval(c.pool, d.n1spline(), p[RT_S_N1VAL], r.lam, e1);
            n2 = fn_eval(c.pool, d.n2spline(), p[RT_S_N2VAL], r.lam, e2);
        }
        const double in2 = rcp(n2);
        mu = n1 * in2;
        if (d.type() == RT_GRATING) L = grating_order(d, r.lam) * r.lam * in2 * p[D_INV_D];
    }
    if constexpr (ITER) {
        if (d.ikind() == IK_CONIC_NEWTON || d.ikind() == IK_TORIC_NEWTON) {
            const NewtonSeq q = taped_iterates(d, t, k0, nit);
            return surface_adjoint(d, entry, nit, mu, L, ad, gp, &q);
        }
    }
    return surface_adjoint(d, entry, nit, mu, L, ad, gp);
}

// bwd_step for kernels reading the derived table: mu = lv[col] and, for a
// grating, L = lv[col + 1], the values the forward step read
template <bool ITER = false, class SD>
__device__ __forceinline__ bool bwd_step_d(const SD& d, const Ray& r, St& ad, const TraceCtx& c, const LaneRows& t,
                                           int k0, double* gp, const double* lv, int col) {
RT_UNROLL
    for (int k = 0; k < PG_N; ++k) gp[k] = 0.0;
    const St entry = {t.ld(k0 + 0), t.ld(k0 + 1), t.ld(k0 + 2), t.ld(k0 + 3), t.ld(k0 + 4), t.ld(k0 + 5)};
    const int nit = (d.ikind() == IK_CONIC_NEWTON || d.ikind() == IK_TORIC_NEWTON) ? (int)t.ld(k0 + 6) : 0;
    double mu = 0.0, L = 0.0;
    if (d.type() == RT_LENS || d.type() == RT_GRATING) {
        mu = lv[col];
        if (d.type() == RT_GRATING) L = lv[col + 1];
    }
    if constexpr (ITER) {
        if (d.ikind() == IK_CONIC_NEWTON || d.ikind() == IK_TORIC_NEWTON) {
            const NewtonSeq q = taped_iterates(d, t, k0, nit);
            return surface_adjoint(d, entry, nit, mu, L, ad, gp, &q);
        }
    }
    return surface_adjoint(d, entry, nit, mu, L, ad, gp);
}

// Gradient accumulators of the backward lambdas: add(j, v) adds the wave's
// sum of v to parameter slot j.  The ray loops call begin/end once per lane
// (adjoint_body) or tile_begin/tile_end per tile (adjoint_batch_body).
//
// WaveAcc: per-wave sums in memory, a.gacc[j * W + wave] (W waves), the
// wave's first lane adding to its own slot (one writer per slot: no race,
// fixed order); slot numbers may be run-time values.
struct WaveAcc {
    LaneRows rows;
    bool lead;
    __device__ void begin(const AdjArgs& a, uint64_t wave, uint64_t W, uint64_t lw) {
        const uint64_t n_par = a.n_par > 0 ? (uint64_t)a.n_par : 1;
        rows = lane_rows(a.gacc, n_par, W, wave);
        lead = lw == 0;
        if (lead)
            for (uint64_t j = 0; j < (uint64_t)a.n_par; ++j) a.gacc[j * W + wave] = 0.0;
    }
    __device__ void end(const AdjArgs&, uint64_t, uint64_t, uint64_t) const {}
    __device__ void add(int j, double v) const {
        const double sum = wave_sum(v);
        if (lead) rows.st(j, rows.ld(j) + sum);
    }
};

// TileOut: the sums of one tile-wave of a batch written to row tw of n_par
// doubles, a.gacc[tw * n_par + j] (tw = tile * waves per workgroup + wave),
// each slot once per tile; slot numbers may be run-time values.
struct TileOut {
    LaneRows rows;                       // row = 8 bytes: slot j at voff + 8 j
    bool lead;
    __device__ void tile_begin(const AdjArgs& a, uint64_t tw, uint64_t lw, uint64_t n_rows) {
        const uint64_t n_par = a.n_par > 0 ? (uint64_t)a.n_par : 1;
        rows.base = a.gacc;
        rows.bytes = (int)(n_rows * n_par * sizeof(double));
        rows.voff = (uint32_t)(tw * n_par * sizeof(double));
        rows.row = (uint32_t)sizeof(double);
        lead = lw == 0;
    }
    __device__ void tile_end(const AdjArgs&, uint64_t, uint64_t) const {}
    __device__ void add(int j, double v) const {
        const double sum = wave_sum(v);
        if (lead) rows.st(j, sum);
    }
};

// GroupAcc: the wave's sums held in registers, for compile-time slot groups
// (the generated kernels).  add4<GRP>(x0, x1, x2, x3) reduces four slots'
// values over the wave together: two 32-lane swaps and one 16-lane swap
// transpose them into the four rows of one register (row 0: x0, row 1: x2,
// row 2: x1, row 3: x3, each summed over 4 lanes), the row sums finish them,
// and lane 16 r + GRP % 16 of v[GRP / 16] accumulates row r.  Map::slot(k)
// names the parameter held by lane k % 64 of v[k / 64] (-1: none).  Written
// once (per lane loop or per tile) to the layouts of WaveAcc and TileOut; no
// memory traffic and no waits inside the ray loop.
template <int NR, class Map>
struct GroupAcc {
    double v[NR];
    uint32_t lw;
    __device__ void zero() {
RT_UNROLL
        for (int q = 0; q < NR; ++q) v[q] = 0.0;
    }
    __device__ void begin(const AdjArgs&, uint64_t, uint64_t, uint64_t lane_in_wave) {
        lw = (uint32_t)lane_in_wave;
        zero();
    }
    __device__ void end(const AdjArgs& a, uint64_t wave, uint64_t W, uint64_t) const {
RT_UNROLL
        for (int q = 0; q < NR; ++q) {
            const int j = Map::slot(q * 64 + (int)lw);
            if (j >= 0) a.gacc[(uint64_t)j * W + wave] = v[q];
        }
    }
    __device__ void tile_begin(const AdjArgs&, uint64_t, uint64_t lane_in_wave, uint64_t) {
        lw = (uint32_t)lane_in_wave;
        zero();
    }
    __device__ void tile_end(const AdjArgs& a, uint64_t tw, uint64_t) const {
RT_UNROLL
        for (int q = 0; q < NR; ++q) {
            const int j = Map::slot(q * 64 + (int)lw);
            if (j >= 0) a.gacc[tw * (uint64_t)a.n_par + (uint64_t)j] = v[q];
        }
    }
    template <int GRP>
    __device__ void add4(double x0, double x1, double x2, double x3) {
        const double t = row_sum(swap_sum<16>(swap_sum<32>(x0, x1), swap_sum<32>(x2, x3)));
        // the lane index made opaque at each use (else the compares with the
        // constant group offsets are hoisted out of the ray loop and held as
        // 64-bit lane masks across it), and the update kept in program order
        // (else the sums are sunk towards the loop end and every surface's
        // partials stay live until there)
        uint32_t l = lw;
        __asm__ volatile("" : "+v"(l));
        double& acc = v[GRP / 16];
        acc += (l & 15u) == (uint32_t)(GRP % 16) ? t : 0.0;
        __asm__ volatile("" : "+v"(acc));
    }
};

// forward(r, alive, c, tape); backward(r, ad, c, tape, acc, ok) -> ok, with
// tape the lane's rows of a.tape (n_surf * kTape rows of G lanes) and acc the
// wave's accumulators in a.gacc (n_par rows of G / 64 waves).  The ray loop is
// wave-uniform (a wave's lanes take rays i0 .. i0 + 63 together and lanes past
// the end run masked), so that the backward steps can reduce across the wave;
// lanes whose ray is lost or whose replay fails compute on and are masked out
// of the sums.  The tape size must fit a buffer resource (checked by the host).
template <int BEAM, class Acc, class Fwd, class Bwd>
__device__ __forceinline__ void adjoint_body(const AdjArgs& a, Fwd forward, Bwd backward) {
    extern __shared__ double s_pool[];
    __shared__ unsigned s_stats[RT_STATS_LEN];
    for (int i = threadIdx.x; i < RT_STATS_LEN; i += blockDim.x) s_stats[i] = 0u;
    for (int i = threadIdx.x; i < a.pool_len; i += blockDim.x) s_pool[i] = a.pool[i];
    __syncthreads();

    TraceCtx c;
    c.pool = s_pool;
    c.st = s_stats;
    c.hist = nullptr;
    c.flux = nullptr;
    c.do_hist = c.do_flux = false;
    c.n_oob = c.n_hit = c.n_off = c.n_edge = 0;
    c.n_prev = 0.0;
    c.last_it = 0;
    const uint64_t G = (uint64_t)gridDim.x * blockDim.x;      // lanes (blockDim a multiple of 64)
    const uint64_t W = G / 64;                                 // waves
    const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t wave = lane / 64, lw = lane % 64;
    Acc acc;
    acc.begin(a, wave, W, lw);
    unsigned skipped = 0;

    for (uint64_t i0 = wave * 64; i0 < a.n; i0 += G) {
        const uint64_t i = i0 + lw;
        const bool in = i < a.n;
        cptr tbl = (cptr)a.table;
        __asm__ volatile("" : "+s"(tbl));
        c.table = tbl;
        Ray r;
        cptr bm = (cptr)__builtin_amdgcn_kernarg_segment_ptr();
        uint64_t jl;
        gen_ray<BEAM>(bm, a.seed, i, r, jl);
        c.lk = (uint32_t)jl;
        bool alive = in;
        forward(r, alive, c, lane_rows(a.tape, a.tape_rows, G, lane));
        unsigned long long good = __ballot(alive), all = __ballot(in);
        if (lw == 0) {
            atomicAdd(&s_stats[RT_ST_RAYS], (unsigned)__popcll(all));
            if (good) atomicAdd(&s_stats[RT_ST_GOOD], (unsigned)__popcll(good));
        }
        if (!good) continue;                                   // wave-uniform
        const uint64_t g = i % a.n_groups;
        const double bx = a.seeds[3 * g], by = a.seeds[3 * g + 1], b = a.seeds[3 * g + 2];
        St ad = {bx + b * r.X, by + b * r.Y, 0.0, 0.0, 0.0, 0.0};
        const bool ok = backward(r, ad, c, lane_rows(a.tape, a.tape_rows, G, lane), acc, alive);
        if (alive && !ok) ++skipped;
    }
    acc.end(a, wave, W, lw);
    lds_add(&s_stats[RT_ST_SPLINE_OOB], c.n_oob);
    lds_add(&s_stats[RT_ST_ADJ_SKIP], skipped);
    __syncthreads();
    if (a.stats) {
        for (int k = threadIdx.x; k < RT_STATS_LEN; k += blockDim.x)
            if (s_stats[k]) atomicAdd(&a.stats[k], (unsigned long long)s_stats[k]);
    }
}

// The batched adjoint: tiles of one instance per workgroup (as the batched
// trace), each lane one ray of the tile; the backward's gradient sums of each
// tile-wave go to a.gacc[(tile * waves_per_group + wave) * n_par + j] and are
// reduced per instance afterwards.  Every wave runs the backward (lanes
// without a live ray masked), so every slot is written.
template <int BEAM, class Acc, class Fwd, class Bwd>
__device__ __forceinline__ void adjoint_batch_body(const AdjArgs& a, Fwd forward, Bwd backward) {
    extern __shared__ double s_pool[];
    __shared__ unsigned s_stats[RT_STATS_LEN];
    for (int i = threadIdx.x; i < RT_STATS_LEN; i += blockDim.x) s_stats[i] = 0u;
    for (int i = threadIdx.x; i < a.pool_len; i += blockDim.x) s_pool[i] = a.pool[i];
    __syncthreads();

    TraceCtx c;
    c.pool = s_pool;
    c.st = s_stats;
    c.hist = nullptr;
    c.flux = nullptr;
    c.do_hist = c.do_flux = false;
    c.n_oob = c.n_hit = c.n_off = c.n_edge = 0;
    c.n_prev = 0.0;
    c.last_it = 0;
    const uint64_t G = (uint64_t)gridDim.x * blockDim.x;      // tape lanes
    const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t lw = threadIdx.x % 64, wpg = blockDim.x / 64, wib = threadIdx.x / 64;
    Acc acc;
    unsigned skipped = 0;

    for (uint64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
        // instance of tile t (uniform).  The batch fields are read through the
        // kernel argument segment
        // (AdjArgs is the kernel's only argument, at offset 0), re-derived
        // here, so that they stay scalar.
        typedef const __attribute__((address_space(4))) AdjArgs* cadj;
        cadj ka = (cadj)__builtin_amdgcn_kernarg_segment_ptr();
        __asm__ volatile("" : "+s"(ka));
        const __attribute__((address_space(4))) uint64_t* ts =
            (const __attribute__((address_space(4))) uint64_t*)ka->tile_start;
        const uint64_t inst = (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)((const __attribute__((address_space(4))) uint32_t*)ka->tile_inst)[t]);
        const __attribute__((address_space(4))) uint64_t* rs =
            (const __attribute__((address_space(4))) uint64_t*)ka->ray_start;
        const uint64_t n_i = rs[inst + 1] - rs[inst];
        const uint64_t r = (t - ts[inst]) * blockDim.x + threadIdx.x;
        const bool in = r < n_i;
        // 32-bit offset (the host keeps n_inst x n_surf records below 2^32
        // doubles); the pointer depends on the tile, so no record load is
        // hoisted out of the tile loop
        cptr tbl = (cptr)ka->table + (uint32_t)inst * (uint32_t)ka->inst_stride;
        c.table = tbl;
        Ray ray;
        cptr bm = ka->inst_beams ? (cptr)ka->inst_beams + inst * (uint64_t)kBeamRec : (cptr)ka;
        uint64_t jl;
        gen_ray<BEAM>(bm, ka->seed + inst, r, ray, jl);
        c.lk = (uint32_t)jl;
        bool alive = in;
        forward(ray, alive, c, lane_rows(a.tape, a.tape_rows, G, lane));
        unsigned long long good = __ballot(alive), all = __ballot(in);
        if (lw == 0) {
            if (all) atomicAdd(&s_stats[RT_ST_RAYS], (unsigned)__popcll(all));
            if (good) atomicAdd(&s_stats[RT_ST_GOOD], (unsigned)__popcll(good));
        }
        const uint64_t g = r % ka->n_groups;
        const double* sd = ka->seeds + (inst * ka->n_groups + g) * 3;
        const double bx = sd[0], by = sd[1], b = sd[2];
        St ad = {bx + b * ray.X, by + b * ray.Y, 0.0, 0.0, 0.0, 0.0};
        acc.tile_begin(a, t * wpg + wib, lw, a.n_tiles * wpg);
        const bool ok = backward(ray, ad, c, lane_rows(a.tape, a.tape_rows, G, lane), acc, alive);
        if (alive && !ok) ++skipped;
        acc.tile_end(a, t * wpg + wib, lw);
    }
    lds_add(&s_stats[RT_ST_SPLINE_OOB], c.n_oob);
    lds_add(&s_stats[RT_ST_ADJ_SKIP], skipped);
    __syncthreads();
    if (a.stats) {
        for (int k = threadIdx.x; k < RT_STATS_LEN; k += blockDim.x)
            if (s_stats[k]) atomicAdd(&a.stats[k], (unsigned long long)s_stats[k]);
    }
}

}  // namespace adj
}  // namespace rt
#endif
