// Mixed-precision forms of the pair and facet force models (HG_PRECISION_MIXED).
//
// Kept in FP64: positions and the centre distance |x_i - x_j| with everything
// derived from it by subtraction (contact overlap, Hamaker gap h, cutoff
// test), the facet closest point, and the spring vectors: their rotation into
// the contact plane, rescaling and the accumulation xi += v dt. In FP32 with
// a*b+c contraction allowed: the normal, relative velocities, normal force,
// Coulomb test and cap, torques and the Hamaker scalar. Dot and cross
// products and a*x+y vector updates, FP64 and FP32, are explicit fused
// multiply-adds (one rounding where the FP64 path has two). The gap h must stay
// FP64: h = |d| - 2r near h0 = 0.1 nm loses eps_32 * 2r/h0 ~ 1.7e-5 of
// relative precision in FP32. Accelerations are summed in FP32 per particle.
//
// The model equations are those of dem_math.h (contact_force.h:57-147,
// alt_sinter_bridge.h:92-156, hamaker_force.h:28-43,
// limited_range_hamaker_force.h:29-47, surface_contact_force.h:57-147,
// surface_hamaker_force.h:18-30); only the arithmetic types differ.
#pragma once

#include "dem_math.h"

namespace hg {

struct f3 {
    float x, y, z;
};
HGD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HGD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HGD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
HGD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
HGD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
HGD float dotf(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HGD f3 crossf(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
HGD f3 tof(d3 a) { return {float(a.x), float(a.y), float(a.z)}; }
HGD d3 tod(f3 a) { return {double(a.x), double(a.y), double(a.z)}; }
HGD f3 zerof() { return {0.f, 0.f, 0.f}; }

// Fused forms for this path: each product folded into the following sum is
// one explicit fused multiply-add, so the rounding is fixed by the source and
// is the same in every kernel that inlines it (alone or packed, chunk or wide,
// pool or single system), and the vector code issues one instruction where a
// multiply and an add were two.
HGD double dotm(d3 a, d3 b) { return __builtin_fma(a.z, b.z, __builtin_fma(a.y, b.y, a.x * b.x)); }
HGD d3 crossm(d3 a, d3 b) {
    return {__builtin_fma(a.y, b.z, -(a.z * b.y)), __builtin_fma(a.z, b.x, -(a.x * b.z)),
            __builtin_fma(a.x, b.y, -(a.y * b.x))};
}
// s a + b
HGD d3 fmam(double s, d3 a, d3 b) {
    return {__builtin_fma(s, a.x, b.x), __builtin_fma(s, a.y, b.y), __builtin_fma(s, a.z, b.z)};
}
HGD float dotmf(f3 a, f3 b) { return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x)); }
HGD f3 crossmf(f3 a, f3 b) {
    return {__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
            __builtin_fmaf(a.x, b.y, -(a.y * b.x))};
}
HGD f3 fmamf(float s, f3 a, f3 b) {
    return {__builtin_fmaf(s, a.x, b.x), __builtin_fmaf(s, a.y, b.y), __builtin_fmaf(s, a.z, b.z)};
}

// Reciprocal square roots without the correctly rounded division/sqrt
// expansions (this path is not bit-exact by design): the hardware estimates
// refined by Newton steps. FP64: v_rsq_f64 then one step
// y <- y (3 - a y^2) / 2, which squares the estimate's relative error (below
// 2^-22) to below 2^-44, 1e-13: a normal or a distance from it is then
// accurate to far below what the FP32 force arithmetic resolves (the torsion
// spring's 1e-5 rad residual, the finest thing derived from the normal, to
// 1e-8 of itself). FP32: v_rsq_f32 / v_rcp_f32 (1 ulp). Arguments must be
// > 0.
__device__ __forceinline__ double rsqrt64(double a) {
#pragma clang fp contract(fast)
    const double y = __builtin_amdgcn_rsq(a);
    return y * (1.5 - 0.5 * a * y * y);
}
__device__ __forceinline__ float rsqrt32(float a) { return __builtin_amdgcn_rsqf(a); }
__device__ __forceinline__ float rcp32(float a) { return __builtin_amdgcn_rcpf(a); }

struct accf {
    f3 lin, ang;
};
HGD accf operator+(accf a, accf b) { return {a.lin + b.lin, a.ang + b.ang}; }

// Inverse masses/inertias and FP32 model constants, prepared once per launch.
struct mixed_consts {
    float c_k, c_gn, c_kt, c_gt, c_mus, c_mud, c_kr, c_gr, c_mur, c_mudr, c_ko, c_go, c_muo, c_mudo, c_r, c_im, c_ii;
    float s_k, s_gn, s_kt, s_gt, s_kr, s_gr, s_ko, s_go, s_r, s_im, s_ii, s_m;
    int s_breaking;
    float h_a6, h_r, h_im, h_2ar;   // -A/6, r, 1/m, 2 A r
    double c_r2, s_r2, h_r2, h_h0, h_cut;
    double c_dt, s_dt;
    float sc_k, sc_gn, sc_kt, sc_gt, sc_mus, sc_mud, sc_kr, sc_gr, sc_mur, sc_mudr, sc_ko, sc_go, sc_muo, sc_mudo, sc_r,
        sc_im, sc_ii;
    double sc_rd, sc_dt;
    float sh_a23, sh_im;
    double sh_r, sh_h0;
    // squared distances: beyond far2 a pair that is not bonded is neither in
    // contact nor within 0.25 r of it (hamaker_far_mx); beyond h_cut2 the
    // limited-range Hamaker force is zero
    double far2, h_cut2;
    float h_r2f;
};

// Holds the FP64 constants of the pair path in vector registers for a pass
// (an empty asm with a vector-register operand): kept as scalars, they are
// spilled and re-read lane by lane at every use inside the pair loop.
__device__ __forceinline__ void pin_vgpr(double &x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ mixed_consts pinned(const mixed_consts &MC) {
    mixed_consts m = MC;
    pin_vgpr(m.far2);
    pin_vgpr(m.h_cut2);
    pin_vgpr(m.c_r2);
    pin_vgpr(m.s_r2);
    pin_vgpr(m.c_dt);
    pin_vgpr(m.s_dt);
    return m;
}

inline mixed_consts make_mixed_consts(const hg_params &p) {
    mixed_consts m;
    const hg_contact_params &c = p.contact;
    m.c_k = float(c.k); m.c_gn = float(c.gamma_n); m.c_kt = float(c.k_t); m.c_gt = float(c.gamma_t);
    m.c_mus = float(c.mu_s); m.c_mud = float(c.mu_s * c.phi_d);
    m.c_kr = float(c.k_r); m.c_gr = float(c.gamma_r); m.c_mur = float(c.mu_r); m.c_mudr = float(c.mu_r * c.phi_r);
    m.c_ko = float(c.k_o); m.c_go = float(c.gamma_o); m.c_muo = float(c.mu_o); m.c_mudo = float(c.mu_o * c.phi_o);
    m.c_r = float(c.r_part); m.c_im = float(1.0 / c.mass); m.c_ii = float(1.0 / c.inertia);
    m.c_r2 = 2.0 * c.r_part; m.c_dt = c.dt;
    const hg_sinter_params &s = p.sinter;
    m.s_k = float(s.k); m.s_gn = float(s.gamma_n); m.s_kt = float(s.k_t); m.s_gt = float(s.gamma_t);
    m.s_kr = float(s.k_r); m.s_gr = float(s.gamma_r); m.s_ko = float(s.k_o); m.s_go = float(s.gamma_o);
    m.s_r = float(s.r_part); m.s_im = float(1.0 / s.mass); m.s_ii = float(1.0 / s.inertia);
    m.s_r2 = 2.0 * s.r_part; m.s_dt = s.dt;
    m.s_m = float(s.mass); m.s_breaking = p.sinter_breaking;
    const hg_hamaker_params &h = p.hamaker;
    m.h_a6 = float(-h.A / 6.0); m.h_r = float(h.r_part); m.h_im = float(1.0 / h.mass);
    m.h_2ar = float(2.0 * h.A * h.r_part);
    m.h_r2 = 2.0 * h.r_part; m.h_h0 = h.h0; m.h_cut = h.r_cutoff;
    const hg_contact_params &sc = p.surface_contact;
    m.sc_k = float(sc.k); m.sc_gn = float(sc.gamma_n); m.sc_kt = float(sc.k_t); m.sc_gt = float(sc.gamma_t);
    m.sc_mus = float(sc.mu_s); m.sc_mud = float(sc.mu_s * sc.phi_d);
    m.sc_kr = float(sc.k_r); m.sc_gr = float(sc.gamma_r); m.sc_mur = float(sc.mu_r); m.sc_mudr = float(sc.mu_r * sc.phi_r);
    m.sc_ko = float(sc.k_o); m.sc_go = float(sc.gamma_o); m.sc_muo = float(sc.mu_o); m.sc_mudo = float(sc.mu_o * sc.phi_o);
    m.sc_r = float(sc.r_part); m.sc_im = float(1.0 / sc.mass); m.sc_ii = float(1.0 / sc.inertia);
    m.sc_rd = sc.r_part; m.sc_dt = sc.dt;
    const hg_hamaker_params &sh = p.surface_hamaker;
    m.sh_a23 = float(2.0 * sh.A / 3.0); m.sh_im = float(1.0 / sh.mass);
    m.sh_r = sh.r_part; m.sh_h0 = sh.h0;
    const double touch = c.r_part > h.r_part ? 2.0 * c.r_part : 2.0 * h.r_part;
    m.far2 = (1.125 * touch) * (1.125 * touch);
    m.h_cut2 = h.r_cutoff * h.r_cutoff;
    m.h_r2f = float(2.0 * h.r_part);
    return m;
}

// Rotates a spring into the plane normal to n and restores its length, in
// FP64 with the FP64 normal. The torsion spring accumulates along n, so its
// projection leaves only the residual from the turn of n since the previous
// step (~|v| dt / r, about 1e-5 rad here), which the rescale then restores to
// full length: the result's direction is that residual's, and an FP32 normal
// (error ~6e-8 rad) would perturb it by ~1%.
// The rescale |xi| / |xr| is x2 / sqrt(x2 l2): one reciprocal square root
// (x2 l2 stays a normal number for springs longer than 1e-77 m).
__device__ __forceinline__ d3 rotate_spring(d3 xi, d3 n) {
#pragma clang fp contract(fast)
    d3 xr = fmam(-dotm(xi, n), n, xi);
    const double l2 = dotm(xr, xr);
    if (l2 > 0.0) {
        const double x2 = dotm(xi, xi);  // > 0 whenever l2 > 0
        xr = xr * (x2 * rsqrt64(x2 * l2));
    }
    return xr;
}

// contact_force.h:105-147 with FP32 forces and an FP64 spring. The spring's
// own updates use the FP64 relative velocity rv64: the static increment
// rv dt and the dynamic reset -(f_d t + gamma rv)/k (with t from the FP64
// test force) set the direction that the next step's projection keeps.
__device__ __forceinline__ f3 shear_contact_mx(d3 &xi, d3 n64, float k, float g, float fn, float mus, float mud,
                                               f3 rv, d3 rv64, double dt) {
#pragma clang fp contract(fast)
    xi = rotate_spring(xi, n64);
    const f3 f0 = fmamf(-k, tof(xi), (-g) * rv);
    const float sf = mus * fn, df = mud * fn;
    const float q = dotmf(f0, f0);
    const float iq = q > 0.f ? rsqrt32(q) : 0.f;
    const float f0n = q * iq;
    const f3 t = f0 * iq;
    float fsel;
    if (f0n <= sf) {
        xi = fmam(dt, rv64, xi);
        fsel = sf;
    } else {
        const double kd = double(k), gd = double(g);
        const d3 f064 = fmam(-kd, xi, (-gd) * rv64);
        const double l2 = dotm(f064, f064);
        const d3 t64 = l2 > 0.0 ? f064 * rsqrt64(l2) : zero3();
        xi = (-1.0 / kd) * fmam(double(df), t64, gd * rv64);
        fsel = df;
    }
    return ((f0n < fsel) ? f0n : fsel) * t;
}

// alt_sinter_bridge.h:137-156 with FP32 forces and an FP64 spring.
__device__ __forceinline__ f3 shear_sinter_mx(d3 &xi, d3 n64, float k, float g, f3 rv, d3 rv64, double dt) {
#pragma clang fp contract(fast)
    xi = rotate_spring(xi, n64);
    const f3 f0 = fmamf(-k, tof(xi), (-g) * rv);
    xi = fmam(dt, rv64, xi);
    return f0;
}

// Shared kinematics of contact and bonded pairs: n from the FP64 distance.
__device__ __forceinline__ void pair_kinematics_mx(d3 xi, d3 xj, double &dist, d3 &n64, f3 &n) {
#pragma clang fp contract(fast)
    const d3 d = xi - xj;
    const double d2 = dotm(d, d);
    // coincident particles (d2 == 0) get a zero normal, as normalized() of a
    // zero vector gives in the reference (Eigen) and in the FP64 path
    const double inv = d2 > 0.0 ? rsqrt64(d2) : 0.0;
    dist = d2 * inv;
    n64 = d * inv;
    n = tof(n64);
}

// The relative-velocity decomposition stays in FP64: when a pair's relative
// velocity is nearly parallel to n, the tangential part
// v_t = v_ij - (v_ij . n) n is a small difference that an FP32 normal would
// leave with an absolute error of eps_32 |v_ij|, seen in full by the torque.
// These are multiplies and adds only; divisions and square roots (normals of
// forces, Coulomb test) are FP32. Identities used:
// n x w_i + n x w_j = n x (w_i + w_j), n x w_i - n x w_j = n x (w_i - w_j).
// vw(vj, wj) gives the partner's velocity and angular velocity, for a bonded
// or overlapping pair only.
template <bool BONDED, class VW>
__device__ __forceinline__ accf bonded_or_contact_mx(const mixed_consts &m, double dist, d3 n64, f3 n, d3 vi, d3 wi,
                                                     VW &&vw, spring9 &sp, bool &active) {
#pragma clang fp contract(fast)
    const double ovd = (BONDED ? m.s_r2 : m.c_r2) - dist;
    if (!BONDED && ovd <= 0.0) {
        active = false;
        sp = zero_spring();
        return {zerof(), zerof()};
    }
    active = true;
    d3 vj, wj;
    vw(vj, wj);
    const float ov = float(ovd);
    const double rd = 0.5 * (BONDED ? m.s_r2 : m.c_r2);
    const double rppd = rd - 0.5 * ovd;
    const float r = float(rd), rpp = float(rppd);
    const d3 dv = vi - vj, dw = wi - wj, sw = wi + wj;
    const float v_n = float(-dotm(dv, n64));
    const float f_n = (BONDED ? m.s_k : m.c_k) * ov + (BONDED ? m.s_gn : m.c_gn) * v_n;
    const d3 v_ij = fmam(rppd, crossm(n64, sw), dv);
    const d3 vt64 = fmam(-dotm(v_ij, n64), n64, v_ij);
    const d3 vr64 = (-rppd) * crossm(n64, dw);
    const d3 vo64 = (rd * dotm(n64, dw)) * n64;
    const f3 v_t = tof(vt64), v_r = tof(vr64), v_o = tof(vo64);
    f3 f_t, f_r, f_o;
    if (BONDED) {
        f_t = shear_sinter_mx(sp.s[0], n64, m.s_kt, m.s_gt, v_t, vt64, m.s_dt);
        f_r = shear_sinter_mx(sp.s[1], n64, m.s_kr, m.s_gr, v_r, vr64, m.s_dt);
        f_o = shear_sinter_mx(sp.s[2], n64, m.s_ko, m.s_go, v_o, vo64, m.s_dt);
    } else {
        f_t = shear_contact_mx(sp.s[0], n64, m.c_kt, m.c_gt, f_n, m.c_mus, m.c_mud, v_t, vt64, m.c_dt);
        f_r = shear_contact_mx(sp.s[1], n64, m.c_kr, m.c_gr, f_n, m.c_mur, m.c_mudr, v_r, vr64, m.c_dt);
        f_o = shear_contact_mx(sp.s[2], n64, m.c_ko, m.c_go, f_n, m.c_muo, m.c_mudo, v_o, vo64, m.c_dt);
    }
    const f3 tau = fmamf(-rpp, crossmf(n, f_t), fmamf(r, crossmf(n, f_r), r * f_o));
    return {fmamf(f_n, n, f_t) * (BONDED ? m.s_im : m.c_im), tau * (BONDED ? m.s_ii : m.c_ii)};
}

// The Hamaker acceleration along n at gap h (h >= h0): the reference's
// direction is normalized(x_j - x_i) = -n.
template <int HAM>
__device__ __forceinline__ f3 hamaker_at_mx(const mixed_consts &m, float h, f3 n) {
#pragma clang fp contract(fast)
    const float rp = m.h_r;
    if constexpr (HAM == HG_HAMAKER_SIMPLIFIED) {
        // simplified_hamaker_force.h:26: 2 A r / h^2 along x_j - x_i = -n
        const float ih = rcp32(h);
        return (-(m.h_2ar * ih * ih) * m.h_im) * n;
    }
    const float a = 4.0f * rp + h, b = 2.0f * rp + h;
    const float ia = rcp32(a), ib = rcp32(b), ih = rcp32(h);
    const float s = m.h_a6 * ((a + h) * ia * ih - 2.0f * ib - 4.0f * rp * rp * ib * ib * ib -
                              2.0f * rp * rp * (a + h) * ia * ia * ih * ih);
    return (-s * m.h_im) * n;
}

// Hamaker (full, limited range or simplified) with an FP64 gap.
template <int HAM>
__device__ __forceinline__ f3 hamaker_mx(const mixed_consts &m, double dist, f3 n) {
#pragma clang fp contract(fast)
    if (HAM == HG_HAMAKER_LIMITED && dist > m.h_cut) return zerof();
    double hd = dist - m.h_r2;
    if (hd < m.h_h0) hd = m.h_h0;
    return hamaker_at_mx<HAM>(m, float(hd), n);
}

// Hamaker of a pair beyond far2 (d2 = |d|^2 > far2, so the gap is at least
// 0.25 r): the distance and normal in FP32 from the FP64 difference d. The
// gap dist - 2r then carries at most eps_32 * 2.25 r / 0.25 r = 5e-7 of
// itself, below what the FP32 force arithmetic resolves anyway.
template <int HAM>
__device__ __forceinline__ f3 hamaker_far_mx(const mixed_consts &m, d3 d, double d2) {
#pragma clang fp contract(fast)
    if (HAM == HG_HAMAKER_LIMITED && d2 > m.h_cut2) return zerof();
    const float q = float(d2);
    const float inv = rsqrt32(q);
    const f3 n = tof(d) * inv;
    return hamaker_at_mx<HAM>(m, q * inv - m.h_r2f, n);
}

__device__ __forceinline__ accf surface_contact_mx(const mixed_consts &m, d3 xf, d3 vf_, d3 xi, d3 vi_, d3 w64,
                                                   spring9 &sp) {
#pragma clang fp contract(fast)
    const d3 d = xi - xf;
    const double d2 = dot(d, d);
    // a centre on its facet point (d2 == 0) gets a zero normal, as normalized()
    // of a zero vector gives in the reference (Eigen) and in the FP64 path
    const double inv = d2 > 0.0 ? rsqrt64(d2) : 0.0;
    const double dist = d2 * inv;
    const double ovd = m.sc_rd - dist;
    if (ovd <= 0.0) {
        sp = zero_spring();
        return {zerof(), zerof()};
    }
    const d3 n64 = d * inv;
    const f3 n = tof(n64);
    const float ov = float(ovd), r = m.sc_r;
    const double rppd = m.sc_rd - ovd;
    const float rpp = float(rppd);
    const d3 dv = vi_ - vf_;
    const float v_n = float(-dot(dv, n64));
    const float f_n = m.sc_k * ov + m.sc_gn * v_n;
    const d3 v_ij = dv + rppd * cross(n64, w64);
    const d3 vt64 = v_ij - dot(v_ij, n64) * n64;
    const d3 vr64 = (-rppd) * cross(n64, w64);
    const d3 vo64 = (m.sc_rd * dot(n64, w64)) * n64;
    const f3 v_t = tof(vt64), v_r = tof(vr64), v_o = tof(vo64);
    const f3 f_t = shear_contact_mx(sp.s[0], n64, m.sc_kt, m.sc_gt, f_n, m.sc_mus, m.sc_mud, v_t, vt64, m.sc_dt);
    const f3 f_r = shear_contact_mx(sp.s[1], n64, m.sc_kr, m.sc_gr, f_n, m.sc_mur, m.sc_mudr, v_r, vr64, m.sc_dt);
    const f3 f_o = shear_contact_mx(sp.s[2], n64, m.sc_ko, m.sc_go, f_n, m.sc_muo, m.sc_mudo, v_o, vo64, m.sc_dt);
    const f3 tau = (-rpp) * crossf(n, f_t) + r * crossf(n, f_r) + r * f_o;
    return {(f_n * n + f_t) * m.sc_im, tau * m.sc_ii};
}

__device__ __forceinline__ f3 surface_hamaker_mx(const mixed_consts &m, d3 xf, d3 xi) {
#pragma clang fp contract(fast)
    const d3 r = xf - xi;
    const double d2 = dot(r, r);
    // a centre on its facet point (d2 == 0) gets a zero normal, as normalized()
    // of a zero vector gives in the reference (Eigen) and in the FP64 path
    const double inv = d2 > 0.0 ? rsqrt64(d2) : 0.0;
    const double dist = d2 * inv;
    double hd = dist - m.sh_r;
    if (hd < m.sh_h0) hd = m.sh_h0;
    const float h = float(hd), rp = float(m.sh_r);
    const float q = h + 2.0f * rp;
    const float s = m.sh_a23 * (rp * rp * rp) * rcp32(h * h * q * q);
    return (s * m.sh_im) * tof(r * inv);
}

}  // namespace hg
