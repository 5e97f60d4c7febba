// Device-side math of the DEM path: vectors, material tables, contact,
// bond and sphere-triangle models.  Each function names the reference
// routine it restates (DEMSPH/kernel/...).
#pragma once

#include "common.hpp"

namespace demsph {
namespace dem {

template <class T>
struct V3 {
    T x, y, z;
};
using D3 = V3<double>;

template <class T>
__host__ __device__ __forceinline__ V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T>
__device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
__device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
__device__ __forceinline__ V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T>
__device__ __forceinline__ V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
__device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
__device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
__device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double len(D3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float len(V3<float> a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// myVec.h:112-128: zero for squared length below 1e-20
__device__ __forceinline__ D3 normalize(D3 a)
{
    const double l2 = a.x * a.x + a.y * a.y + a.z * a.z;
    if (l2 < 1.0e-20) return {0.0, 0.0, 0.0};
    const double inv = 1.0 / sqrt(l2);
    return {a.x * inv, a.y * inv, a.z * inv};
}
template <class T, class S>
__device__ __forceinline__ V3<T> cvt(V3<S> a) { return {(T)a.x, (T)a.y, (T)a.z}; }
// myVec.h:46-50
__device__ __forceinline__ D3 divide(D3 a, double s)
{
    const double inv = 1.0 / s;
    return {a.x * inv, a.y * inv, a.z * inv};
}

// Rodrigues rotation of v by rotation vector a (myVec.h:134-176).
__device__ __forceinline__ D3 rotate(D3 v, D3 a)
{
    const double t2 = dot(a, a);
    if (t2 < 1.0e-20) return v;
    const double t = sqrt(t2);
    double s, c;
    sincos(t, &s, &c);
    const double inv = 1.0 / t;
    const D3 k = a * inv;
    const double kv = dot(k, v);
    const D3 kxv = cross(k, v);
    const double s3 = kv * (1.0 - c);
    return {v.x * c + kxv.x * s + k.x * s3, v.y * c + kxv.y * s + k.y * s3, v.z * c + kxv.z * s + k.z * s3};
}

struct Quat {
    double q0, q1, q2, q3;
};
// myQua.h:48-62, 77-92
__device__ __forceinline__ Quat qintegrate(Quat q, D3 w, double dt)
{
    const D3 v = w * (0.5 * dt);
    Quat d;
    d.q0 = -q.q1 * v.x - q.q2 * v.y - q.q3 * v.z;
    d.q1 = q.q0 * v.x + q.q3 * v.y - q.q2 * v.z;
    d.q2 = -q.q3 * v.x + q.q0 * v.y + q.q1 * v.z;
    d.q3 = q.q2 * v.x - q.q1 * v.y + q.q0 * v.z;
    Quat s{q.q0 + d.q0, q.q1 + d.q1, q.q2 + d.q2, q.q3 + d.q3};
    const double l2 = s.q0 * s.q0 + s.q1 * s.q1 + s.q2 * s.q2 + s.q3 * s.q3;
    if (l2 < 1.0e-20) return s;
    const double inv = 1.0 / sqrt(l2);
    return {s.q0 * inv, s.q1 * inv, s.q2 * inv, s.q3 * inv};
}
// myQua.h:103-113
__device__ __forceinline__ D3 qrotate(Quat q, D3 v)
{
    const D3 qv{q.q1, q.q2, q.q3};
    const D3 t = cross(qv, v) * 2.0;
    return v + t * q.q0 + cross(qv, t);
}

struct Sym {
    double xx, yy, zz, xy, xz, yz;
};
// R * invI * R^T (myMat.h:107-158)
__device__ __forceinline__ Sym sym_rotate(Quat q, Sym m)
{
    const double q0 = q.q0, q1 = q.q1, q2 = q.q2, q3 = q.q3;
    const double q11 = 2 * q1 * q1, q22 = 2 * q2 * q2, q33 = 2 * q3 * q3;
    const double a = 1.0 - q22 - q33, b = 2 * (q1 * q2 - q0 * q3), c = 2 * (q1 * q3 + q0 * q2);
    const double d = 2 * (q1 * q2 + q0 * q3), e = 1.0 - q11 - q33, f = 2 * (q2 * q3 - q0 * q1);
    const double g = 2 * (q1 * q3 - q0 * q2), h = 2 * (q2 * q3 + q0 * q1), i = 1.0 - q11 - q22;
    const double r0x = m.xx * a + m.xy * b + m.xz * c, r0y = m.xy * a + m.yy * b + m.yz * c,
                 r0z = m.xz * a + m.yz * b + m.zz * c;
    const double r1x = m.xx * d + m.xy * e + m.xz * f, r1y = m.xy * d + m.yy * e + m.yz * f,
                 r1z = m.xz * d + m.yz * e + m.zz * f;
    const double r2x = m.xx * g + m.xy * h + m.xz * i, r2y = m.xy * g + m.yy * h + m.yz * i,
                 r2z = m.xz * g + m.yz * h + m.zz * i;
    Sym r;
    r.xx = r0x * a + r0y * b + r0z * c;
    r.yy = r1x * d + r1y * e + r1z * f;
    r.zz = r2x * g + r2y * h + r2z * i;
    r.xy = r0x * d + r0y * e + r0z * f;
    r.xz = r0x * g + r0y * h + r0z * i;
    r.yz = r1x * g + r1y * h + r1z * i;
    return r;
}
__device__ __forceinline__ D3 sym_mul(Sym m, D3 v)
{
    return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
            m.xz * v.x + m.yz * v.y + m.zz * v.z};
}

// ------------------------------------------------------------ material table
// Packed like contactModelParameters (contactKernel.cu:6-104): param-major
// arrays of cap = n(n+1)/2 + 1 slots, the last one the fallback for
// material ids out of range (contactParameters.h:123-141).
enum { H_E = 0, H_G, H_RES, H_KRKS, H_KTKS, H_MUS, H_MUR, H_MUT, H_D, H_N };
enum { L_KN = 0, L_KS, L_KR, L_KT, L_DN, L_DS, L_DR, L_DT, L_MUS, L_MUR, L_MUT, L_N };
enum { B_GAMMA = 0, B_E, B_KNKS, B_SIG, B_C, B_MU, B_N };

struct Tables {
    int nMat, cap;
    const double* H;  // H_N x cap (H_D: derived -ln e / sqrt(ln^2 e + pi^2))
    const double* L;  // L_N x cap
    const double* B;  // B_N x cap
};

__host__ __device__ __forceinline__ int pair_index(int a, int b, int nMat, int cap)
{
    if (nMat <= 0 || cap <= 0) return 0;
    if (a < 0 || b < 0 || a >= nMat || b >= nMat) return cap - 1;
    int i = a, j = b;
    if (i > j) { int t = i; i = j; j = t; }
    long long idx = ((long long)i * (2LL * nMat - i + 1LL)) / 2LL + (long long)(j - i);
    if (idx < 0) idx = 0;
    if (idx >= cap) idx = cap - 1;
    return (int)idx;
}

__device__ __forceinline__ double Lp(const Tables& t, int ip, int q) { return t.cap ? t.L[q * t.cap + ip] : 0.0; }
__device__ __forceinline__ double Hp(const Tables& t, int ip, int q)
{
    return t.cap ? t.H[q * t.cap + ip] : (q == H_RES ? 1.0 : 0.0);
}
__device__ __forceinline__ double Bp(const Tables& t, int ip, int q)
{
    return t.cap ? t.B[q * t.cap + ip] : ((q == B_GAMMA || q == B_KNKS) ? 1.0 : 0.0);
}

// ------------------------------------------------------------ contact model
// integrateSlidingOrRollingSpring (contactKernel.cuh:64-97).  The spring
// (a tangential displacement) is FP64; the force arithmetic is in T.
template <class T>
__device__ __forceinline__ D3 spring_slide(D3 prev, V3<T> sv, V3<T> n, V3<T> Fn, T mu, T k, T c, double dt)
{
    D3 s{0.0, 0.0, 0.0};
    if (k > (T)0) {
        const D3 nd = cvt<double>(n);
        D3 p1 = prev - nd * dot(prev, nd);
        const double a = len(p1);
        if (a > 1.e-10) p1 = p1 * (len(prev) / a);
        s = p1 + cvt<double>(sv) * dt;
        if (mu > (T)0) {
            V3<T> f = cvt<T>(s) * (-k) - sv * c;
            const T af = len(f), an = len(Fn);
            if (af > mu * an) {
                const T ratio = mu * an / af;
                f = f * ratio;
                s = cvt<double>((f + sv * c) * (-(T)1 / k));
            }
        }
    }
    return s;
}

// integrateTorsionSpring (contactKernel.cuh:99-126)
template <class T>
__device__ __forceinline__ D3 spring_torsion(D3 prev, V3<T> tv, V3<T> n, V3<T> Fn, T mu, T k, T c, double dt)
{
    D3 s{0.0, 0.0, 0.0};
    if (k > (T)0) {
        const D3 nd = cvt<double>(n);
        s = nd * dot(prev + cvt<double>(tv) * dt, nd);
        if (mu > (T)0) {
            V3<T> f = cvt<T>(s) * (-k) - tv * c;
            const T af = len(f), an = len(Fn);
            if (af > mu * an) {
                const T ratio = mu * an / af;
                f = f * ratio;
                s = cvt<double>((f + tv * c) * (-(T)1 / k));
            }
        }
    }
    return s;
}

// LinearContact / HertzianMindlinContact (contactKernel.cuh:128-290) with
// the model choice of contactKernel.cu:218-252 (linear if k_n > 1e-20).
// delta <= 0 zeroes the springs and leaves F, Tq, Fs untouched.  Fs is the
// tangential (sliding) part of F: the lever arm of a contact is parallel to
// n, so only Fs carries a moment about a ball centre.
template <class T>
__device__ __forceinline__ void contact_model(const Tables& tb, int ip, double m, double R, double delta, V3<T> vc,
                                              V3<T> wij, V3<T> n, double dt, D3& es, D3& er, D3& et, V3<T>& F,
                                              V3<T>& Tq, V3<T>& Ftan)
{
    if (!(delta > 0)) {
        es = er = et = D3{0.0, 0.0, 0.0};
        return;
    }
    T kn, ks, kr, kt, cn, cs, cr, ct, mus, mur, mut;
    if (Lp(tb, ip, L_KN) > 1.e-20) {
        kn = (T)Lp(tb, ip, L_KN); ks = (T)Lp(tb, ip, L_KS); kr = (T)Lp(tb, ip, L_KR); kt = (T)Lp(tb, ip, L_KT);
        const T mm = (T)m;
        cn = (T)2 * (T)Lp(tb, ip, L_DN) * sqrt(mm * kn);
        cs = (T)2 * (T)Lp(tb, ip, L_DS) * sqrt(mm * ks);
        cr = (T)2 * (T)Lp(tb, ip, L_DR) * sqrt(mm * kr);
        ct = (T)2 * (T)Lp(tb, ip, L_DT) * sqrt(mm * kt);
        mus = (T)Lp(tb, ip, L_MUS); mur = (T)Lp(tb, ip, L_MUR); mut = (T)Lp(tb, ip, L_MUT);
    } else {
        const T D = (T)Hp(tb, ip, H_D);
        const T sq = (T)sqrt(R * delta);
        kn = (T)(4. / 3.) * (T)Hp(tb, ip, H_E) * sq;
        ks = (T)8 * (T)Hp(tb, ip, H_G) * sq;
        const T mm = (T)m;
        const T c56 = (T)(2.0 * 0.91287092917527685576);  // 2 sqrt(5/6)
        cn = c56 * D * sqrt(mm * kn);
        cs = c56 * D * sqrt(mm * ks);
        kr = ks * (T)Hp(tb, ip, H_KRKS);
        kt = ks * (T)Hp(tb, ip, H_KTKS);
        cr = c56 * D * sqrt(mm * kr);
        ct = c56 * D * sqrt(mm * kt);
        mus = (T)Hp(tb, ip, H_MUS); mur = (T)Hp(tb, ip, H_MUR); mut = (T)Hp(tb, ip, H_MUT);
    }
    const T Rt = (T)R;
    const V3<T> vn = n * dot(vc, n);
    const V3<T> Fn = n * (kn * (T)delta) - vn * cn;
    const V3<T> vs = vc - vn;
    es = spring_slide<T>(es, vs, n, Fn, mus, ks, cs, dt);
    const V3<T> Fs = cvt<T>(es) * (-ks) - vs * cs;
    const V3<T> vr = cross(n, wij) * (-Rt);
    er = spring_slide<T>(er, vr, n, Fn, mur, kr, cr, dt);
    const V3<T> Fr = cvt<T>(er) * (-kr) - vr * cr;
    const V3<T> Tr = cross(n, Fr) * Rt;
    const T D2 = (T)2 * Rt;
    const V3<T> vt = n * (D2 * dot(wij, n));
    et = spring_torsion<T>(et, vt, n, Fn, mut, kt, ct, dt);
    const V3<T> Ft = cvt<T>(et) * (-kt) - vt * ct;
    F = Fn + Fs;
    Ftan = Fs;
    Tq = Tr + Ft * D2;
}

// ParallelBondedContact (contactKernel.cuh:4-62), FP64: the bond state is
// accumulated incrementally over the whole run.
__device__ __forceinline__ int bond_model(double& Fn, double& Tt, D3& Fs, D3& Tb, double& sig, double& tau, D3 n0,
                                          D3 n, D3 vc, D3 wa, D3 wb, double ra, double rb, double dt, double gam,
                                          double E, double knks, double sigt, double C, double mu)
{
    const D3 th1 = cross(n0, n);
    Fs = rotate(Fs, th1);
    Tb = rotate(Tb, th1);
    const D3 th2 = n * dot((wa + wb) * 0.5 * dt, n);
    Fs = rotate(Fs, th2);
    Tb = rotate(Tb, th2);
    const double rmin = ra < rb ? ra : rb;
    const double Rb = gam * rmin;
    const double A = Rb * Rb * M_PI;
    const double I = Rb * Rb * Rb * Rb / 4. * M_PI;
    const double J = 2 * I;
    const double kn = E / (ra + rb);
    const double ks = kn / knks;
    const D3 dn = n * (dot(vc, n) * dt);
    const D3 dtg = vc * dt - dn;
    Fn -= dot(dn * kn * A, n);
    Fs = Fs - dtg * ks * A;
    const D3 wr = wa - wb;
    const D3 rn = n * (dot(wr, n) * dt);
    const D3 rt = wr * dt - rn;
    Tt -= dot(rn * ks * J, n);
    Tb = Tb - rt * kn * I;
    sig = -Fn / A + len(Tb) / I * Rb;
    tau = len(Fs) / A + fabs(Tt) / J * Rb;
    int bonded = 1;
    if (sigt > 0 && sig > sigt) bonded = 0;
    else if (C > 0 && tau > C - mu * sig) bonded = 0;
    return bonded;
}

// ------------------------------------------------------------ sphere-triangle
enum { ST_NONE = 0, ST_FACE = 1, ST_EDGE = 2, ST_VERTEX = 3 };

__device__ __forceinline__ double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }
__device__ __forceinline__ double dist2(D3 a, D3 b)
{
    const D3 t = a - b;
    return dot(t, t);
}
// closestPointOnSegment (contactKernel.cuh:315-330)
__device__ __forceinline__ D3 seg_closest(D3 a, D3 b, D3 p, double& t)
{
    const D3 ab = b - a;
    const double ab2 = dot(ab, ab);
    if (ab2 <= 1e-30) {
        t = 0.0;
        return a;
    }
    t = clamp01(dot(p - a, ab) / ab2);
    return a + ab * t;
}

// classifySphereTriangleContact (contactKernel.cuh:367-501): Ericson's
// region tests (Real-Time Collision Detection, 5.1.5); degenerate
// triangles as three segments.
__device__ __forceinline__ int classify(D3 c, double r, D3 v0, D3 v1, D3 v2, D3& cp)
{
    const double r2 = r * r;
    const double eps2 = 1e-12 * fmax(1.0, r2);
    const D3 ab = v1 - v0, ac = v2 - v0;
    const D3 n = cross(ab, ac);
    const double area2 = dot(n, n);
    if (area2 <= 1e-20) {
        double t01, t02, t12;
        const D3 q01 = seg_closest(v0, v1, c, t01), q02 = seg_closest(v0, v2, c, t02), q12 = seg_closest(v1, v2, c, t12);
        double dmin = dist2(c, q01);
        cp = q01;
        int type = (t01 > 0.0 && t01 < 1.0) ? ST_EDGE : ST_VERTEX;
        const double d02 = dist2(c, q02);
        if (d02 < dmin) { dmin = d02; cp = q02; type = (t02 > 0.0 && t02 < 1.0) ? ST_EDGE : ST_VERTEX; }
        const double d12 = dist2(c, q12);
        if (d12 < dmin) { dmin = d12; cp = q12; type = (t12 > 0.0 && t12 < 1.0) ? ST_EDGE : ST_VERTEX; }
        return (dmin <= r2 + eps2) ? type : ST_NONE;
    }
    const D3 ap = c - v0;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { cp = v0; return dist2(c, v0) <= r2 + eps2 ? ST_VERTEX : ST_NONE; }
    const D3 bp = c - v1;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { cp = v1; return dist2(c, v1) <= r2 + eps2 ? ST_VERTEX : ST_NONE; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double t = d1 / (d1 - d3);
        cp = v0 + ab * t;
        return dist2(c, cp) <= r2 + eps2 ? ST_EDGE : ST_NONE;
    }
    const D3 cp2 = c - v2;
    const double d5 = dot(ab, cp2), d6 = dot(ac, cp2);
    if (d6 >= 0.0 && d5 <= d6) { cp = v2; return dist2(c, v2) <= r2 + eps2 ? ST_VERTEX : ST_NONE; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double t = d2 / (d2 - d6);
        cp = v0 + ac * t;
        return dist2(c, cp) <= r2 + eps2 ? ST_EDGE : ST_NONE;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        cp = v1 + (v2 - v1) * t;
        return dist2(c, cp) <= r2 + eps2 ? ST_EDGE : ST_NONE;
    }
    const double sum = va + vb + vc;
    if (fabs(sum) <= 1e-30) return ST_NONE;
    const double inv = 1.0 / sum;
    const double v = vb * inv, w = vc * inv;
    cp = v0 + ab * v + ac * w;
    return dist2(c, cp) <= r2 + eps2 ? ST_FACE : ST_NONE;
}

// triangleCircumcenter (wallIntegrationKernel.cuh:4-29)
__host__ __device__ __forceinline__ void circumcenter(const double* a, const double* b, const double* c, double* out)
{
    const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
    const double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (n2 < 1e-30) {
        const double inv3 = 1.0 / 3.0;
        for (int k = 0; k < 3; k++) out[k] = (a[k] + b[k] + c[k]) * inv3;
        return;
    }
    const double nab[3] = {n[1] * ab[2] - n[2] * ab[1], n[2] * ab[0] - n[0] * ab[2], n[0] * ab[1] - n[1] * ab[0]};
    const double acn[3] = {ac[1] * n[2] - ac[2] * n[1], ac[2] * n[0] - ac[0] * n[2], ac[0] * n[1] - ac[1] * n[0]};
    const double lac = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
    const double lab = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    const double inv = 1.0 / (2.0 * n2);
    for (int k = 0; k < 3; k++) out[k] = a[k] + (nab[k] * lac + acn[k] * lab) * inv;
}

}  // namespace dem
}  // namespace demsph
