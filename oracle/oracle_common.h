/*
 * Internal helpers of the FP64 CPU oracle (test infrastructure only).
 * Vector algebra restated from the reference's utility header
 * (DEMSPH/kernel/myUtility/myVec.h:25-176, myQua.h, myMat.h): same
 * operation order, double precision, host sqrt.
 */
#ifndef DEMSPH_ORACLE_COMMON_H
#define DEMSPH_ORACLE_COMMON_H

#include <math.h>
#include "oracle.h"

#define O_PI 3.14159265358979323846

static inline o_vec3 v3(double x, double y, double z) { o_vec3 r = {x, y, z}; return r; }
static inline o_vec3 vld(const double* p, long i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void vst(double* p, long i, o_vec3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
static inline o_vec3 vadd(o_vec3 a, o_vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline o_vec3 vsub(o_vec3 a, o_vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline o_vec3 vmul(o_vec3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline o_vec3 vneg(o_vec3 a) { return v3(-a.x, -a.y, -a.z); }
/* myVec.h:46-50: division through the reciprocal */
static inline o_vec3 vdiv(o_vec3 a, double s) { double inv = 1.0 / s; return v3(a.x * inv, a.y * inv, a.z * inv); }
static inline double vdot(o_vec3 a, o_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline o_vec3 vcross(o_vec3 a, o_vec3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vlen2(o_vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double vlen(o_vec3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* myVec.h:112-128: zero below 1e-20 squared length */
static inline o_vec3 vnormalize(o_vec3 a)
{
    double l2 = a.x * a.x + a.y * a.y + a.z * a.z;
    if (l2 < 1.0e-20) return v3(0.0, 0.0, 0.0);
    double inv = 1.0 / sqrt(l2);
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
/* Rodrigues rotation of v by the rotation vector a (myVec.h:134-176). */
static inline o_vec3 vrotate(o_vec3 v, o_vec3 a)
{
    double t2 = vdot(a, a);
    if (t2 < 1.0e-20) return v;
    double t = sqrt(t2);
    double s = sin(t), c = cos(t);
    double inv = 1.0 / t;
    o_vec3 k = v3(a.x * inv, a.y * inv, a.z * inv);
    double kv = k.x * v.x + k.y * v.y + k.z * v.z;
    o_vec3 kxv = v3(k.y * v.z - k.z * v.y, k.z * v.x - k.x * v.z, k.x * v.y - k.y * v.x);
    double s3 = kv * (1.0 - c);
    return v3(v.x * c + kxv.x * s + k.x * s3, v.y * c + kxv.y * s + k.y * s3, v.z * c + kxv.z * s + k.z * s3);
}

typedef struct { double q0, q1, q2, q3; } o_quat;
typedef struct { double xx, yy, zz, xy, xz, yz; } o_sym;

static inline o_quat qnormalize(o_quat q)
{
    double l2 = q.q0 * q.q0 + q.q1 * q.q1 + q.q2 * q.q2 + q.q3 * q.q3;
    if (l2 < 1.0e-20) return q;
    double inv = 1.0 / sqrt(l2);
    o_quat r = {q.q0 * inv, q.q1 * inv, q.q2 * inv, q.q3 * inv};
    return r;
}
/* q <- normalize(q + 0.5*q*(0,w*dt)) (myQua.h:77-92). */
static inline o_quat qintegrate(o_quat q, o_vec3 w, double dt)
{
    o_vec3 v = vmul(w, 0.5 * dt);
    o_quat d;
    d.q0 = -q.q1 * v.x - q.q2 * v.y - q.q3 * v.z;
    d.q1 = q.q0 * v.x + q.q3 * v.y - q.q2 * v.z;
    d.q2 = -q.q3 * v.x + q.q0 * v.y + q.q1 * v.z;
    d.q3 = q.q2 * v.x - q.q1 * v.y + q.q0 * v.z;
    o_quat s = {q.q0 + d.q0, q.q1 + d.q1, q.q2 + d.q2, q.q3 + d.q3};
    return qnormalize(s);
}
/* v' = v + q0*t + qv x t, t = 2 qv x v (myQua.h:103-113). */
static inline o_vec3 qrotate(o_quat q, o_vec3 v)
{
    o_vec3 qv = v3(q.q1, q.q2, q.q3);
    o_vec3 t = vmul(vcross(qv, v), 2.0);
    return vadd(vadd(v, vmul(t, q.q0)), vcross(qv, t));
}
/* R * invI * R^T (myMat.h:107-158). */
static inline o_sym sym_rotate(o_quat q, o_sym m)
{
    double q0 = q.q0, q1 = q.q1, q2 = q.q2, q3 = q.q3;
    double q11 = 2 * q1 * q1, q22 = 2 * q2 * q2, q33 = 2 * q3 * q3;
    double a = 1.0 - q22 - q33, b = 2 * (q1 * q2 - q0 * q3), c = 2 * (q1 * q3 + q0 * q2);
    double d = 2 * (q1 * q2 + q0 * q3), e = 1.0 - q11 - q33, f = 2 * (q2 * q3 - q0 * q1);
    double g = 2 * (q1 * q3 - q0 * q2), h = 2 * (q2 * q3 + q0 * q1), i = 1.0 - q11 - q22;
    double r0x = m.xx * a + m.xy * b + m.xz * c, r0y = m.xy * a + m.yy * b + m.yz * c, r0z = m.xz * a + m.yz * b + m.zz * c;
    double r1x = m.xx * d + m.xy * e + m.xz * f, r1y = m.xy * d + m.yy * e + m.yz * f, r1z = m.xz * d + m.yz * e + m.zz * f;
    double r2x = m.xx * g + m.xy * h + m.xz * i, r2y = m.xy * g + m.yy * h + m.yz * i, r2z = m.xz * g + m.yz * h + m.zz * i;
    o_sym r;
    r.xx = r0x * a + r0y * b + r0z * c;
    r.yy = r1x * d + r1y * e + r1z * f;
    r.zz = r2x * g + r2y * h + r2z * i;
    r.xy = r0x * d + r0y * e + r0z * f;
    r.xz = r0x * g + r0y * h + r0z * i;
    r.yz = r1x * g + r1y * h + r1z * i;
    return r;
}
static inline o_vec3 sym_mulv(o_sym m, o_vec3 v)
{
    return v3(m.xx * v.x + m.xy * v.y + m.xz * v.z, m.xy * v.x + m.yy * v.y + m.yz * v.z,
              m.xz * v.x + m.yz * v.y + m.zz * v.z);
}
/* Inverse of a symmetric 3x3 matrix by cofactors (myMat.h:163-179). */
static inline o_sym sym_inverse(o_sym A)
{
    double det_inv = 1.0 / (A.xx * (A.yy * A.zz - A.yz * A.yz) + A.xy * (A.yz * A.xz - A.xy * A.zz) +
                            A.xz * (A.xy * A.yz - A.yy * A.xz));
    o_sym r;
    r.xx = (A.yy * A.zz - A.yz * A.yz) * det_inv;
    r.yy = (A.xx * A.zz - A.xz * A.xz) * det_inv;
    r.zz = (A.xx * A.yy - A.xy * A.xy) * det_inv;
    r.xy = -(A.xy * A.zz - A.xz * A.yz) * det_inv;
    r.xz = (A.xy * A.yz - A.xz * A.yy) * det_inv;
    r.yz = -(A.xx * A.yz - A.xz * A.xy) * det_inv;
    return r;
}

static inline int o_hash_cell(const o_grid* g, int x, int y, int z) { return z * g->ny * g->nx + y * g->nx + x; }

#endif
