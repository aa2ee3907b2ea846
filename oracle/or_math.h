/*
 * ORACLE (test infrastructure only).  CPU restatement of the reference's
 * math layer, templated on the reference's `real` (float = USE_FLOAT build,
 * double = FP64 build).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline.
 *
 * Follows (paths relative to the reference's src/engine/):
 *   include/engine/common.h:18-56        real, PI, eps, eps_pdf
 *   include/engine/core/utils.h:11-197   sampling and vector helpers
 *   src/math/transform.cpp:9-63          Transform (TRS, inverse, apply)
 * glm semantics restated from its public documentation: normalize(v) =
 * v * (1/sqrt(dot(v,v))), column-major mat4, lookAt under
 * GLM_FORCE_LEFT_HANDED (common.h:4).
 */
#ifndef FM_ORACLE_MATH_H
#define FM_ORACLE_MATH_H

#include <cmath>
#include <cstdint>
#include <limits>
#include <algorithm>

namespace orc {

template <typename R> struct V2 {
    R x, y;
    V2() : x(0), y(0) {}
    V2(R a, R b) : x(a), y(b) {}
};
template <typename R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    explicit V3(R s) : x(s), y(s), z(s) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <typename R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename R> inline V3<R>& operator+=(V3<R>& a, V3<R> b) { a = a + b; return a; }
template <typename R> inline V3<R>& operator*=(V3<R>& a, V3<R> b) { a = a * b; return a; }
template <typename R> inline V3<R>& operator*=(V3<R>& a, R s) { a = a * s; return a; }
template <typename R> inline V3<R>& operator/=(V3<R>& a, R s) { a = a / s; return a; }
template <typename R> inline V2<R> operator+(V2<R> a, V2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> inline V2<R> operator*(R s, V2<R> a) { return {s * a.x, s * a.y}; }

/* glm::dot sums the component products left to right. */
template <typename R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <typename R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <typename R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / std::sqrt(dot(a, a))); }

template <typename R> struct K {
    /* common.h:50-56 */
    static constexpr R REAL_MAX = std::numeric_limits<R>::max();
    static constexpr R REAL_MIN = std::numeric_limits<R>::min();
    static R PI() { return (R)3.14159265354; }
    static R InvPI() { return (R)(1.0 / (double)PI()); }
    static R eps() { return (R)1e-4; }
    static R eps_pdf() { return (R)1e-20; }
};

/* utils.h helpers */
template <typename R> inline R Clamp(R v, R lo, R hi) { return v > hi ? hi : (v < lo ? lo : v); }
template <typename R> inline R AbsDot(V3<R> a, V3<R> b) { return std::abs(dot(a, b)); }
template <typename R> inline R PowerHeuristic(R f, R g) { return (f * f) / (f * f + g * g); }
template <typename R> inline R SquareSum(V3<R> v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
template <typename R> inline R Luminance(V3<R> c) { return (R)0.2126 * c.x + (R)0.7152 * c.y + (R)0.0722 * c.z; }
template <typename R> inline bool IsBlack(V3<R> c) {
    if (c.x == R(0) && c.y == R(0) && c.z == R(0)) return true;
    if (c.x < R(0) || c.y < R(0) || c.z < R(0)) return true;
    return false;
}
template <typename R> inline R DegToRad(R deg) { return deg / (R)180 * K<R>::PI(); }
template <typename R> inline R SrgbToLinear(R c) {
    if (c <= (R)0.04045) return c * R(1) / (R)12.92;
    return std::pow((c + (R)0.055) * R(1) / (R)1.055, (R)2.4);
}
template <typename R> inline V2<R> UniformDisk(V2<R> s) {
    R r = std::sqrt(s.x);
    R th = R(2) * K<R>::PI() * s.y;
    return {r * std::cos(th), r * std::sin(th)};
}
template <typename R> inline V3<R> CosineWeightedHemiSphere(V2<R> s) {
    V2<R> p = UniformDisk(s);
    R zz = R(1) - p.x * p.x - p.y * p.y;
    if (zz < R(0)) zz = R(0);
    return normalize(V3<R>(p.x, p.y, std::sqrt(zz)));
}
template <typename R> inline V3<R> UniformSphere(V2<R> s) {
    R z = R(1) - R(2) * s.x;
    R r = std::sqrt(std::max(R(0), R(1) - z * z));
    R phi = R(2) * K<R>::PI() * s.y;
    return {r * std::cos(phi), r * std::sin(phi), z};
}
template <typename R> inline void UniformCone(R cos_t_min, V2<R> s, V3<R>* pos, R* pdf) {
    const R cos_t = cos_t_min + s.x * (R(1) - cos_t_min);
    const R sin_t = std::sqrt(std::max(R(0), R(1) - cos_t * cos_t));
    const R phi = R(2) * K<R>::PI() * s.y;
    pos->x = std::cos(phi) * sin_t;
    pos->y = std::sin(phi) * sin_t;
    pos->z = cos_t;
    *pdf = R(1) / (R(2) * K<R>::PI() * (R(1) - cos_t_min));
}
template <typename R> inline void CreateCoordSys(V3<R> v1, V3<R>* v2, V3<R>* v3) {
    if (std::abs(v1.x) > std::abs(v1.y)) *v2 = normalize(V3<R>(-v1.z, R(0), v1.x));
    else *v2 = normalize(V3<R>(R(0), v1.z, -v1.y));
    *v3 = cross(v1, *v2);
}
template <typename R> inline R SphericalTheta(V3<R> d) { return std::acos(Clamp(d.z, R(-1), R(1))); }
template <typename R> inline R SphericalPhi(V3<R> d) {
    const R p = std::atan2(d.y, d.x);
    return (p < 0) ? (p + R(2) * K<R>::PI()) : p;
}
template <typename R> inline V3<R> SphericalUV(V3<R> d) {
    const R th = SphericalTheta(d);
    const R ph = SphericalPhi(d);
    return {ph * K<R>::InvPI() * (R)0.5, th * K<R>::InvPI(), th};
}
template <typename R> inline V3<R> ToTint(V3<R> c) {
    const R l = Luminance(c);
    return l > R(0) ? c * (R(1) / l) : V3<R>(R(1));
}
template <typename R> inline R CosDir(V3<R> w) { return normalize(w).z; }
template <typename R> inline R AbsCosDir(V3<R> w) { return std::abs(CosDir(w)); }
/* std::min / std::max keep the first argument unless the second compares
 * strictly smaller / larger: this fixes their NaN behaviour (bbox.h:59-69). */
template <typename R> inline R smin(R a, R b) { return (b < a) ? b : a; }
template <typename R> inline R smax(R a, R b) { return (a < b) ? b : a; }
template <typename R> inline V3<R> MinVec(V3<R> a, V3<R> b) { return {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
template <typename R> inline V3<R> MaxVec(V3<R> a, V3<R> b) { return {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
template <typename R> inline R MaxElem(V3<R> v) { return v.x > v.y ? (v.x > v.z ? v.x : v.z) : (v.y > v.z ? v.y : v.z); }
template <typename R> inline R MinElem(V3<R> v) { return v.x < v.y ? (v.x < v.z ? v.x : v.z) : (v.y < v.z ? v.y : v.z); }
template <typename R> inline V3<R> Lerp3(V3<R> a, V3<R> b, R t) { return (R(1) - t) * a + t * b; }
template <typename R> inline R Lerp1(R a, R b, R t) { return (R(1) - t) * a + t * b; }

/* Column-major 4x4 (glm layout: m[col][row]). */
template <typename R> struct M4 {
    R m[4][4];
    static M4 ident() { M4 r; for (int c = 0; c < 4; c++) for (int k = 0; k < 4; k++) r.m[c][k] = (c == k) ? R(1) : R(0); return r; }
};
template <typename R> inline M4<R> mul(const M4<R>& a, const M4<R>& b) {
    M4<R> r;
    for (int c = 0; c < 4; c++)
        for (int k = 0; k < 4; k++)
            r.m[c][k] = a.m[0][k] * b.m[c][0] + a.m[1][k] * b.m[c][1] + a.m[2][k] * b.m[c][2] + a.m[3][k] * b.m[c][3];
    return r;
}
/* (m * v) component k = sum_c m[c][k] v[c] */
template <typename R> inline void mulv(const M4<R>& a, const R v[4], R out[4]) {
    for (int k = 0; k < 4; k++)
        out[k] = a.m[0][k] * v[0] + a.m[1][k] * v[1] + a.m[2][k] * v[2] + a.m[3][k] * v[3];
}
template <typename R> inline M4<R> transpose(const M4<R>& a) {
    M4<R> r;
    for (int c = 0; c < 4; c++) for (int k = 0; k < 4; k++) r.m[c][k] = a.m[k][c];
    return r;
}
/* General inverse by cofactors (adjugate / determinant). */
template <typename R> inline M4<R> inverse(const M4<R>& A) {
    /* work on row-major a[r][c] = A.m[c][r] */
    R a[4][4];
    for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) a[r][c] = A.m[c][r];
    R inv[4][4];
    auto minor3 = [&](int skip_r, int skip_c) {
        R s[3][3]; int rr = 0;
        for (int r = 0; r < 4; r++) {
            if (r == skip_r) continue;
            int cc = 0;
            for (int c = 0; c < 4; c++) { if (c == skip_c) continue; s[rr][cc++] = a[r][c]; }
            rr++;
        }
        return s[0][0] * (s[1][1] * s[2][2] - s[1][2] * s[2][1])
             - s[0][1] * (s[1][0] * s[2][2] - s[1][2] * s[2][0])
             + s[0][2] * (s[1][0] * s[2][1] - s[1][1] * s[2][0]);
    };
    R cof[4][4];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            R mnr = minor3(r, c);
            cof[r][c] = ((r + c) & 1) ? -mnr : mnr;
        }
    R det = a[0][0] * cof[0][0] + a[0][1] * cof[0][1] + a[0][2] * cof[0][2] + a[0][3] * cof[0][3];
    R id = R(1) / det;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) inv[r][c] = cof[c][r] * id;  /* adjugate = cofactor^T */
    M4<R> out;
    for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) out.m[c][r] = inv[r][c];
    return out;
}
/* glm::rotate(angle, axis): Rodrigues rotation about normalize(axis). */
template <typename R> inline M4<R> rotate(R ang, V3<R> axis_in) {
    const R c = std::cos(ang), s = std::sin(ang);
    V3<R> ax = normalize(axis_in);
    V3<R> t = (R(1) - c) * ax;
    M4<R> r = M4<R>::ident();
    r.m[0][0] = c + t.x * ax.x;  r.m[0][1] = t.x * ax.y + s * ax.z; r.m[0][2] = t.x * ax.z - s * ax.y;
    r.m[1][0] = t.y * ax.x - s * ax.z; r.m[1][1] = c + t.y * ax.y;  r.m[1][2] = t.y * ax.z + s * ax.x;
    r.m[2][0] = t.z * ax.x + s * ax.y; r.m[2][1] = t.z * ax.y - s * ax.x; r.m[2][2] = c + t.z * ax.z;
    return r;
}
/* glm::lookAt under GLM_FORCE_LEFT_HANDED (lookAtLH). */
template <typename R> inline M4<R> lookAtLH(V3<R> eye, V3<R> center, V3<R> up) {
    V3<R> f = normalize(center - eye);
    V3<R> s = normalize(cross(up, f));
    V3<R> u = cross(f, s);
    M4<R> r = M4<R>::ident();
    r.m[0][0] = s.x; r.m[1][0] = s.y; r.m[2][0] = s.z;
    r.m[0][1] = u.x; r.m[1][1] = u.y; r.m[2][1] = u.z;
    r.m[0][2] = f.x; r.m[1][2] = f.y; r.m[2][2] = f.z;
    r.m[3][0] = -dot(s, eye); r.m[3][1] = -dot(u, eye); r.m[3][2] = -dot(f, eye);
    return r;
}

template <typename R> struct Ray {
    /* core/ray.h:9-36 */
    V3<R> ori, dir;
    R t_max, t_min;
    Ray() : ori(), dir(R(1), R(0), R(0)), t_max(K<R>::REAL_MAX), t_min(R(0)) {}
    Ray(V3<R> o, V3<R> d, R tmin = R(0), R tmax = K<R>::REAL_MAX) : ori(o), dir(d), t_max(tmax), t_min(tmin) {}
    V3<R> at(R t) const { return ori + dir * t; }
    bool Contains(R t) const { return t >= t_min && t <= t_max; }
};

template <typename R> struct HitPoint;

/* src/math/transform.cpp:9-63 */
template <typename R> struct Transform {
    M4<R> mat, inv;
    Transform() : mat(M4<R>::ident()), inv(M4<R>::ident()) {}
    Transform(const M4<R>& m, const M4<R>& i) : mat(m), inv(i) {}
    explicit Transform(const M4<R>& m) : mat(m), inv(inverse(m)) {}
    static Transform TRS(V3<R> tr, V3<R> rot, V3<R> sc) {
        M4<R> rz = rotate(DegToRad(rot.z), V3<R>(R(0), R(0), R(1)));
        M4<R> ry = rotate(DegToRad(rot.y), V3<R>(R(0), R(1), R(0)));
        M4<R> rx = rotate(DegToRad(rot.x), V3<R>(R(1), R(0), R(0)));
        M4<R> r = mul(mul(rz, ry), rx);
        M4<R> t = M4<R>::ident(); t.m[3][0] = tr.x; t.m[3][1] = tr.y; t.m[3][2] = tr.z;
        M4<R> s = M4<R>::ident(); s.m[0][0] = sc.x; s.m[1][1] = sc.y; s.m[2][2] = sc.z;
        M4<R> m = mul(mul(t, r), s);
        return Transform(m, inverse(m));
    }
    Transform Inv() const { return Transform(inv, mat); }
    V3<R> Vec(V3<R> v) const { R a[4] = {v.x, v.y, v.z, R(0)}, o[4]; mulv(mat, a, o); return {o[0], o[1], o[2]}; }
    V3<R> Point(V3<R> v) const { R a[4] = {v.x, v.y, v.z, R(1)}, o[4]; mulv(mat, a, o); return {o[0] / o[3], o[1] / o[3], o[2] / o[3]}; }
    V3<R> Normal(V3<R> n) const {
        M4<R> it = transpose(inv);
        R a[4] = {n.x, n.y, n.z, R(0)}, o[4]; mulv(it, a, o);
        return normalize(V3<R>(o[0], o[1], o[2]));
    }
    Ray<R> ApplyRay(const Ray<R>& r) const { return Ray<R>(Point(r.ori), Vec(r.dir), r.t_min, r.t_max); }
};

}  // namespace orc
#endif
