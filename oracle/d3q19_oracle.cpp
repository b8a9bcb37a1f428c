// ORACLE — test infrastructure only.
//
// A CPU restatement of Palabos's D3Q19 collide-and-stream path, written to
// follow the reference's arithmetic and its operation order term by term, so
// that in FP64 (and FP32) it reproduces compiled Palabos bit for bit when both
// are compiled without floating-point contraction. Only tests/, the smoke check
// and the benchmark's CPU-baseline leg may load it, and only as the checker.
// The product library never links or calls it.
//
// Streaming is done with two arrays (pull), not with Palabos's in-place swap;
// the cell update order does not change any value (each cell's collision reads
// only its own populations), so the results are identical.
//
// Parity status of each function (see the tests named in brackets):
//   d3q19 constants, moments, BGK, RBGK, streaming + edge reflection,
//   periodicity, BounceBack, NoDynamics, MomentumExchangeBounceBack,
//   Zou-He velocity/pressure, regularized velocity/pressure
//        -> pinned bit-exact against compiled Palabos golden vectors
//           [tests/test_oracle_golden.py]
//   FP16 / BF16 storage rounding
//        -> pinned against numpy's float16 conversion and against exhaustive
//           bit-pattern checks [tests/test_oracle_invariants.py]
//   BGK + Guo body force
//        -> pinned bit-exact against compiled Palabos (ForcedD3Q19 +
//           GuoExternalForceBGKdynamics) golden vectors
//   RBGK + Guo, Brinkman porosity term (no Palabos counterpart)
//        -> pinned per cell by the closed forms of the post-collision mass,
//           first and second moments and of the porous statistic, at random
//           rho, 0 < alpha < 1 and F != 0, each shown to reject the plausible
//           mistakes (regularizing around j instead of j_eq, damping j instead
//           of u, the wrong velocity in the Guo source, PiNeq from j_eq)
//           [tests/test_oracle_invariants.py:
//            test_forced_porous_collision_moments_closed_form,
//            test_rbgk_guo_without_porosity_closed_form,
//            test_brinkman_intermediate_alpha_without_force];
//           alpha = 0 and F = 0 reduce to the Palabos-pinned paths, alpha = 1
//           cells relax to rest [test_brinkman_solid_relaxes_momentum_geometrically]

#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>

namespace {

// D3Q19 velocity set in Palabos order: rest, nine "negative-first" velocities,
// then their opposites (i <-> i+9).
// latticeBoltzmann/nearestNeighborLattices3D.hh:166-178
const int C[19][3] = {
    { 0, 0, 0},
    {-1, 0, 0}, { 0,-1, 0}, { 0, 0,-1},
    {-1,-1, 0}, {-1, 1, 0}, {-1, 0,-1},
    {-1, 0, 1}, { 0,-1,-1}, { 0,-1, 1},
    { 1, 0, 0}, { 0, 1, 0}, { 0, 0, 1},
    { 1, 1, 0}, { 1,-1, 0}, { 1, 0, 1},
    { 1, 0,-1}, { 0, 1, 1}, { 0, 1,-1}
};

inline int opposite(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }  // indexTemplates.h:42-46

template<typename T> struct W {
    // nearestNeighborLattices3D.hh:186-199
    static T t(int i) {
        if (i == 0) return (T)1 / (T)3;
        int n2 = C[i][0]*C[i][0] + C[i][1]*C[i][1] + C[i][2]*C[i][2];
        return n2 == 1 ? (T)1 / (T)18 : (T)1 / (T)36;
    }
    static T cs2() { return (T)1 / (T)3; }
};

// ---------------------------------------------------------------------------
// Storage conversions (round to nearest even).

inline uint16_t f32_to_f16(float f) {
    uint32_t x; std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) {  // inf or nan
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    }
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520 -> inf
    int e = (int)(ax >> 23) - 127;
    uint32_t mant = ax & 0x7fffffu;
    if (e >= -14) {  // normal half
        uint32_t m = mant | 0x800000u;           // 24-bit significand
        uint32_t shift = 13;
        uint32_t q = m >> shift, r = m & ((1u << shift) - 1), half = 1u << (shift - 1);
        if (r > half || (r == half && (q & 1u))) q += 1;
        // q in [2^10, 2^11]; carry to 2^11 bumps the exponent
        uint32_t he = (uint32_t)(e + 15);
        if (q == 0x800u) { q = 0x400u; he += 1; }
        if (he >= 31) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | (he << 10) | (q & 0x3ffu));
    }
    // subnormal half: value = q * 2^-24
    if (e < -25) return (uint16_t)sign;  // below half of the smallest subnormal -> 0 (ties handled below)
    uint32_t m = mant | 0x800000u;           // value = m * 2^(e-23)
    uint32_t shift = (uint32_t)(-e - 1);     // q = m * 2^(e-23+24) = m >> (-1-e)
    uint32_t q = m >> shift;
    uint32_t r = m & ((1u << shift) - 1), half = 1u << (shift - 1);
    if (r > half || (r == half && (q & 1u))) q += 1;
    return (uint16_t)(sign | q);             // q == 0x400 becomes the smallest normal, as it should
}

inline float f16_to_f32(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) x = sign;
        else {
            float v = (float)m * 5.9604644775390625e-08f;  // m * 2^-24, exact
            std::memcpy(&x, &v, 4);
            x |= sign;
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f; std::memcpy(&f, &x, 4);
    return f;
}

inline uint16_t f32_to_bf16(float f) {
    uint32_t x; std::memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);  // quiet nan
    uint32_t lsb = (x >> 16) & 1u;
    x += 0x7fffu + lsb;
    return (uint16_t)(x >> 16);
}

inline float bf16_to_f32(uint16_t h) {
    uint32_t x = (uint32_t)h << 16;
    float f; std::memcpy(&f, &x, 4);
    return f;
}

}  // namespace

extern "C" {

// Layouts shared with the Python side (tests/oracle.py builds these with ctypes).
typedef struct {
    int32_t family;     // 0 Zou-He, 1 regularized
    int32_t kind;       // 0 velocity, 1 pressure
    int32_t dir, ori;   // plane normal axis, orientation (-1 or +1)
    int32_t box[6];     // x0 x1 y0 y1 z0 z1, inclusive
    const double* values;  // velocity: 3 per box cell; pressure: 1 (rho) per box cell; x,y,z C order
    int32_t shape;      // 0 plane, 1 external velocity edge (dir = edge axis), 2 external velocity corner
    int32_t normal[3];  // edges / corners: outward normal (0 along an edge's axis)
    int32_t bc_type;    // boundary::BcType (boundaryCondition.h:40): 0 dirichlet, 1 neumann,
                        // 2 freeslip, 3 density, 4 outflow, 5 normalOutflow
} OraclePlane;

typedef struct {
    int32_t nx, ny, nz;
    int32_t storage;     // 0 f64, 1 f32, 2 f16 storage / f32 compute, 3 bf16 storage / f32 compute
    int32_t dynamics;    // 0 BGK, 1 regularized BGK
    int32_t periodic[3];
    double omega;
    const uint8_t* tags; // nullable; per cell: 0 bulk, 1 BounceBack, 2 NoDynamics, 3 MomentumExchangeBounceBack
    int32_t nplanes;
    const OraclePlane* planes;
    double force[3];     // Guo body force per unit volume (all zero: plain dynamics)
    const double* alpha; // nullable; Brinkman porosity per cell in [0,1]
    int32_t bc_wrapped;  // wrapped boundary factories: processors read plane cells completed
} OracleDesc;

}  // extern "C"

namespace {

enum { TAG_BULK = 0, TAG_BB = 1, TAG_NODYN = 2, TAG_MEBB = 3, TAG_WALL = 4, TAG_PLANE = 5, TAG_STORE = 6 };

// Moments ------------------------------------------------------------------

// momentTemplates3D.h:351-373 (partial_rho) and :411-423 (get_rhoBar_j)
template<typename T>
inline void get_rhoBar_j(const T* f, T& rhoBar, T j[3]) {
    T sXM1 = f[1] + f[4] + f[5] + f[6] + f[7];
    T sXP1 = f[10] + f[13] + f[14] + f[15] + f[16];
    T sYM1 = f[2] + f[4] + f[8] + f[9] + f[14];
    T sYP1 = f[5] + f[11] + f[13] + f[17] + f[18];
    T sZM1 = f[3] + f[6] + f[8] + f[16] + f[18];
    T sZP1 = f[7] + f[9] + f[12] + f[15] + f[17];
    T sX0 = f[0] + f[2] + f[3] + f[8] + f[9] + f[11] + f[12] + f[17] + f[18];
    rhoBar = sXM1 + sX0 + sXP1;
    j[0] = (sXP1 - sXM1);
    j[1] = (sYP1 - sYM1);
    j[2] = (sZP1 - sZM1);
}

template<typename T> inline T invRhoOf(T rhoBar) { return (T)1 / (rhoBar + (T)1); }  // roundOffPolicy.h:86-88

// momentTemplates3D.h:517-539 (compute_rhoBar_j_PiNeq, D3Q19); PiNeq order xx xy xz yy yz zz
template<typename T>
inline void get_rhoBar_j_PiNeq(const T* f, T& rhoBar, T j[3], T pi[6]) {
    T sXM1 = f[1] + f[4] + f[5] + f[6] + f[7];
    T sXP1 = f[10] + f[13] + f[14] + f[15] + f[16];
    T sYM1 = f[2] + f[4] + f[8] + f[9] + f[14];
    T sYP1 = f[5] + f[11] + f[13] + f[17] + f[18];
    T sZM1 = f[3] + f[6] + f[8] + f[16] + f[18];
    T sZP1 = f[7] + f[9] + f[12] + f[15] + f[17];
    T sX0 = f[0] + f[2] + f[3] + f[8] + f[9] + f[11] + f[12] + f[17] + f[18];
    rhoBar = sXM1 + sX0 + sXP1;
    j[0] = (sXP1 - sXM1);
    j[1] = (sYP1 - sYM1);
    j[2] = (sZP1 - sZM1);
    T invRho = invRhoOf(rhoBar);
    T cs2 = W<T>::cs2();
    pi[0] = sXP1 + sXM1 - cs2*rhoBar - invRho*j[0]*j[0];
    pi[3] = sYP1 + sYM1 - cs2*rhoBar - invRho*j[1]*j[1];
    pi[5] = sZP1 + sZM1 - cs2*rhoBar - invRho*j[2]*j[2];
    pi[1] = f[4] - f[5] + f[13] - f[14] - invRho*j[0]*j[1];
    pi[2] = f[6] - f[7] + f[15] - f[16] - invRho*j[0]*j[2];
    pi[4] = f[8] - f[9] + f[17] - f[18] - invRho*j[1]*j[2];
}

// dynamicsTemplates3D.h:1994-1997 (bgk_ma2_equilibrium, D3Q19)
template<typename T>
inline T equilibrium(int i, T rhoBar, T invRho, const T j[3], T jSqr) {
    T c_j = C[i][0]*j[0] + C[i][1]*j[1] + C[i][2]*j[2];
    return W<T>::t(i) * (rhoBar + (T)3.*c_j + invRho*((T)4.5*c_j*c_j - (T)1.5*jSqr));
}

// momentTemplates3D.h:375-382 (get_rhoBar, D3Q19: plain sum in index order)
template<typename T>
inline T get_rhoBar(const T* f) {
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8]
         + f[9] + f[10] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
}

// externalForceTemplates3D.h:73-150 (addGuoForce, ForcedD3Q19, amplitude 1)
template<typename T>
inline void addGuoForce(T* f, const T u[3], const T force[3], T omega) {
    T mu = (T)1*((T)1-omega/(T)2);
    const T oneOver6 = (T)1/(T)6;
    const T oneOver12 = (T)1/(T)12;
    f[0] += -mu*(force[0]*u[0]+force[1]*u[1]+force[2]*u[2]);
    f[1] += oneOver6*mu*(force[0]*(-(T)1+2*u[0])-force[1]*u[1]-force[2]*u[2]);
    f[2] += -oneOver6*mu*(force[0]*u[0]+force[1]*((T)1-2*u[1])+force[2]*u[2]);
    f[3] += -oneOver6*mu*(force[0]*u[0] + force[1]*u[1] + force[2]*((T)1-2*u[2]));
    f[4] += oneOver12*mu*(force[0]*(-(T)1+2*u[0]+3*u[1]) + force[1]*(-(T)1+2*u[1]+3*u[0]) - force[2]*u[2]);
    f[5] += oneOver12*mu*(force[0]*(-(T)1+2*u[0]-3*u[1]) + force[1]*((T)1+2*u[1]-3*u[0]) - force[2]*u[2]);
    f[6] += oneOver12*mu*(force[0]*(-(T)1+2*u[0]+3*u[2]) - force[1]*u[1] + force[2]*(-(T)1+2*u[2]+3*u[0]));
    f[7] += oneOver12*mu*(force[0]*(-(T)1+2*u[0]-3*u[2]) - force[1]*u[1] + force[2]*((T)1+2*u[2]-3*u[0]));
    f[8] += -oneOver12*mu*(force[0]*u[0] + force[1]*((T)1-2*u[1]-3*u[2]) + force[2]*((T)1-2*u[2]-3*u[1]));
    f[9] += -oneOver12*mu*(force[0]*u[0] + force[1]*((T)1-2*u[1]+3*u[2]) + force[2]*(-(T)1-2*u[2]+3*u[1]));
    f[10] += oneOver6*mu*(force[0]*((T)1+2*u[0]) -force[1]*u[1] -force[2]*u[2]);
    f[11] += -oneOver6*mu*(force[0]*u[0] +force[1]*(-(T)1-2*u[1]) +force[2]*u[2]);
    f[12] += -oneOver6*mu*(force[0]*u[0] +force[1]*u[1] +force[2]*(-(T)1-2*u[2]));
    f[13] += oneOver12*mu*(force[0]*((T)1+2*u[0]+3*u[1]) +force[1]*((T)1+2*u[1]+3*u[0]) -force[2]*u[2]);
    f[14] += oneOver12*mu*(force[0]*((T)1+2*u[0]-3*u[1]) +force[1]*(-(T)1+2*u[1]-3*u[0]) -force[2]*u[2]);
    f[15] += oneOver12*mu*(force[0]*((T)1+2*u[0]+3*u[2]) -force[1]*u[1] +force[2]*((T)1+2*u[2]+3*u[0]));
    f[16] += oneOver12*mu*(force[0]*((T)1+2*u[0]-3*u[2]) -force[1]*u[1] +force[2]*(-(T)1+2*u[2]-3*u[0]));
    f[17] += -oneOver12*mu*(force[0]*u[0] +force[1]*(-(T)1-2*u[1]-3*u[2]) +force[2]*(-(T)1-2*u[2]-3*u[1]));
    f[18] += -oneOver12*mu*(force[0]*u[0] +force[1]*(-(T)1-2*u[1]+3*u[2]) +force[2]*((T)1-2*u[2]+3*u[1]));
}

// momentTemplates3D.h:482-496 (compute_PiNeq about a given rhoBar and j, with
// invRho(rhoBar)): the stress the regularized collideExternal uses.
template<typename T>
inline void compute_PiNeq_given(const T* f, T rhoBar, const T j[3], T pi[6]) {
    T sXM1 = f[1] + f[4] + f[5] + f[6] + f[7];
    T sXP1 = f[10] + f[13] + f[14] + f[15] + f[16];
    T sYM1 = f[2] + f[4] + f[8] + f[9] + f[14];
    T sYP1 = f[5] + f[11] + f[13] + f[17] + f[18];
    T sZM1 = f[3] + f[6] + f[8] + f[16] + f[18];
    T sZP1 = f[7] + f[9] + f[12] + f[15] + f[17];
    T invRho = invRhoOf(rhoBar);
    T cs2 = W<T>::cs2();
    pi[0] = sXP1 + sXM1 - cs2*rhoBar - invRho*j[0]*j[0];
    pi[3] = sYP1 + sYM1 - cs2*rhoBar - invRho*j[1]*j[1];
    pi[5] = sZP1 + sZM1 - cs2*rhoBar - invRho*j[2]*j[2];
    pi[1] = f[4] - f[5] + f[13] - f[14] - invRho*j[0]*j[1];
    pi[2] = f[6] - f[7] + f[15] - f[16] - invRho*j[0]*j[2];
    pi[4] = f[8] - f[9] + f[17] - f[18] - invRho*j[1]*j[2];
}

// Collisions ---------------------------------------------------------------

// The constants of the BGK update f_i <- one_m_omega f_i + t_omega (C1 + C2 + C3),
// and the relaxation rate they stand for.
template<typename T>
struct Rates { T one_m_omega, t0_omega, t1_omega, t4_omega, omega; };

// Palabos's (dynamicsTemplates3D.h:2291-2294).
template<typename T>
inline Rates<T> palabosRates(T omega) {
    return {(T)1 - omega, W<T>::t(0) * omega, W<T>::t(1) * omega, W<T>::t(4) * omega, omega};
}

// The 16-bit storage modes (no Palabos counterpart): omega' = 36 w with
// w = rint(omega 2^25 / 36) 2^-25 (the product in FP64); 12w, 2w, w and
// 1 - 36w are then exact, so the update returns rhoBar and j with the factor
// one_m_omega + 36 w = 1 exactly, where Palabos's rounded t_k omega leave a
// constant bias of order 1e-7 per step.
inline Rates<float> exactRates(float omega) {
    const float q = (float)std::nearbyint((double)omega * (33554432.0 / 36.0));
    const float w = q * 0x1p-25f;
    const float om = 36.0f * w;
    return {1.0f - om, 12.0f * w, 2.0f * w, w, om};
}

// dynamicsTemplates3D.h:2289-2368 (bgk_ma2_collision_base, D3Q19), with the rates `r`
template<typename T>
inline T bgk_collision(T* f, T rhoBar, const T j[3], const Rates<T>& r, T invRho) {
    T one_m_omega = r.one_m_omega;
    T t0_omega = r.t0_omega;
    T t1_omega = r.t1_omega;
    T t4_omega = r.t4_omega;
    T jSqr = j[0]*j[0] + j[1]*j[1] + j[2]*j[2];
    T kx = (T)3 * j[0];
    T ky = (T)3 * j[1];
    T kz = (T)3 * j[2];
    T kxSqr_ = invRho / (T)2 * kx*kx;
    T kySqr_ = invRho / (T)2 * ky*ky;
    T kzSqr_ = invRho / (T)2 * kz*kz;
    T kxky_ = invRho * kx*ky;
    T kxkz_ = invRho * kx*kz;
    T kykz_ = invRho * ky*kz;
    T C1 = rhoBar + invRho*(T)3*jSqr;
    T C2, C3;
    C3 = -kxSqr_ - kySqr_ - kzSqr_;
    f[0] *= one_m_omega; f[0] += t0_omega * (C1+C3);
    C2 = -kx; C3 = -kySqr_ - kzSqr_;
    f[1] *= one_m_omega; f[1] += t1_omega * (C1+C2+C3);
    f[10] *= one_m_omega; f[10] += t1_omega * (C1-C2+C3);
    C2 = -ky; C3 = -kxSqr_ - kzSqr_;
    f[2] *= one_m_omega; f[2] += t1_omega * (C1+C2+C3);
    f[11] *= one_m_omega; f[11] += t1_omega * (C1-C2+C3);
    C2 = -kz; C3 = -kxSqr_ - kySqr_;
    f[3] *= one_m_omega; f[3] += t1_omega * (C1+C2+C3);
    f[12] *= one_m_omega; f[12] += t1_omega * (C1-C2+C3);
    C2 = -kx - ky; C3 = kxky_ - kzSqr_;
    f[4] *= one_m_omega; f[4] += t4_omega * (C1+C2+C3);
    f[13] *= one_m_omega; f[13] += t4_omega * (C1-C2+C3);
    C2 = -kx + ky; C3 = -kxky_ - kzSqr_;
    f[5] *= one_m_omega; f[5] += t4_omega * (C1+C2+C3);
    f[14] *= one_m_omega; f[14] += t4_omega * (C1-C2+C3);
    C2 = -kx - kz; C3 = kxkz_ - kySqr_;
    f[6] *= one_m_omega; f[6] += t4_omega * (C1+C2+C3);
    f[15] *= one_m_omega; f[15] += t4_omega * (C1-C2+C3);
    C2 = -kx + kz; C3 = -kxkz_ - kySqr_;
    f[7] *= one_m_omega; f[7] += t4_omega * (C1+C2+C3);
    f[16] *= one_m_omega; f[16] += t4_omega * (C1-C2+C3);
    C2 = -ky - kz; C3 = kykz_ - kxSqr_;
    f[8] *= one_m_omega; f[8] += t4_omega * (C1+C2+C3);
    f[17] *= one_m_omega; f[17] += t4_omega * (C1-C2+C3);
    C2 = -ky + kz; C3 = -kykz_ - kxSqr_;
    f[9] *= one_m_omega; f[9] += t4_omega * (C1+C2+C3);
    f[18] *= one_m_omega; f[18] += t4_omega * (C1-C2+C3);
    return invRho*invRho*jSqr;
}

// dynamicsTemplates3D.h:1902-1973 (neqPiD3Q19::fromPiToFneq0..9); pi order xx xy xz yy yz zz
template<typename T>
inline void neqPi(const T pi[6], T p[10]) {
    const T xx = pi[0], xy = pi[1], xz = pi[2], yy = pi[3], yz = pi[4], zz = pi[5];
    p[0] = (T)3./(T)2 * (-(T)1./(T)3*xx - (T)1./(T)3*yy - (T)1./(T)3*zz);
    p[1] = (T)1./(T)4 * ((T)2./(T)3*xx - (T)1./(T)3*yy - (T)1./(T)3*zz);
    p[2] = (T)1./(T)4 * (-(T)1./(T)3*xx + (T)2./(T)3*yy - (T)1./(T)3*zz);
    p[3] = (T)1./(T)4 * (-(T)1./(T)3*xx - (T)1./(T)3*yy + (T)2./(T)3*zz);
    p[4] = (T)1./(T)8 * ((T)2./(T)3*xx + (T)2./(T)3*yy - (T)1./(T)3*zz + (T)2*xy);
    p[5] = (T)1./(T)8 * ((T)2./(T)3*xx + (T)2./(T)3*yy - (T)1./(T)3*zz - (T)2*xy);
    p[6] = (T)1./(T)8 * ((T)2./(T)3*xx - (T)1./(T)3*yy + (T)2./(T)3*zz + (T)2*xz);
    p[7] = (T)1./(T)8 * ((T)2./(T)3*xx - (T)1./(T)3*yy + (T)2./(T)3*zz - (T)2*xz);
    p[8] = (T)1./(T)8 * (-(T)1./(T)3*xx + (T)2./(T)3*yy + (T)2./(T)3*zz + (T)2*yz);
    p[9] = (T)1./(T)8 * (-(T)1./(T)3*xx + (T)2./(T)3*yy + (T)2./(T)3*zz - (T)2*yz);
}

// dynamicsTemplates3D.h:2500-2554 (rlb_collision, D3Q19)
template<typename T>
inline T rlb_collision(T* f, T rhoBar, T invRho, const T j[3], const T pi[6], T omega) {
    const T jSqr = j[0]*j[0] + j[1]*j[1] + j[2]*j[2];
    T p[10];
    neqPi(pi, p);
    for (int i = 0; i < 10; ++i)
        f[i] = equilibrium(i, rhoBar, invRho, j, jSqr) + ((T)1-omega) * p[i];
    for (int i = 10; i < 19; ++i)
        f[i] = equilibrium(i, rhoBar, invRho, j, jSqr) + ((T)1-omega) * p[i-9];
    return jSqr*invRho*invRho;
}

// offEquilibriumTemplates.h:91-113 (generic fromPiToFneq, used by regularize)
template<typename T>
inline T fromPiToFneq(int i, const T pi[6]) {
    T fNeq = T();
    int iPi = 0;
    for (int a = 0; a < 3; ++a) {
        fNeq += pi[iPi] * (C[i][a]*C[i][a] - W<T>::cs2());
        ++iPi;
        for (int b = a + 1; b < 3; ++b) {
            fNeq += pi[iPi] * (T)2 * C[i][a]*C[i][b];
            ++iPi;
        }
    }
    fNeq *= W<T>::t(i) * (T)3 * (T)3 / (T)2;
    return fNeq;
}

// ---------------------------------------------------------------------------
// The lattice being stepped: compute type T, storage type S.

struct Stats {
    double sumRhoBar = 0, sumUSqr = 0, maxUSqr = 0;
    long count = 0;
    double force[3] = {0, 0, 0};
    double porous[3] = {0, 0, 0};
};

// Storage conversion of one population (defined below, per storage type).
template<typename S, typename T> struct Store;

template<typename T>
struct Engine {
    const OracleDesc& d;
    long N;
    std::vector<uint8_t> tag;          // TAG_*
    std::vector<int> planeOf;          // per cell, -1 or plane index
    std::vector<long> boxIndexOf;      // per cell, index in the plane box
    std::vector<uint32_t> fluidDirs;   // per cell, bit i set: direction i is a fluid direction (MEBB)
    const T* alphaT = nullptr;         // porosity in the compute type (derivative-carrying runs)
    T planeScale = T(1.0);             // scale of every velocity plane's imposed velocity
    // Stored boundary values of every plane, edge and corner (the velocity or
    // density their dynamics hold; the Neumann-type processors rewrite them).
    mutable std::vector<std::vector<double>> vals;

    explicit Engine(const OracleDesc& d_) : d(d_) {
        N = (long)d.nx * d.ny * d.nz;
        tag.assign(N, TAG_BULK);
        planeOf.assign(N, -1);
        boxIndexOf.assign(N, 0);
        fluidDirs.assign(N, 0);
        if (d.tags) for (long k = 0; k < N; ++k) tag[k] = d.tags[k];
        vals.resize(d.nplanes);
        for (int p = 0; p < d.nplanes; ++p) {
            const int* b = d.planes[p].box;
            const long cells = (long)(b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1);
            const long per = (d.planes[p].shape == 0 && d.planes[p].kind == 1) ? 1 : 3;
            vals[p].assign(d.planes[p].values, d.planes[p].values + cells * per);
            long idx = 0;
            for (int x = b[0]; x <= b[1]; ++x)
                for (int y = b[2]; y <= b[3]; ++y)
                    for (int z = b[4]; z <= b[5]; ++z, ++idx) {
                        long k = lin(x, y, z);
                        tag[k] = d.planes[p].shape == 0 ? TAG_PLANE : TAG_STORE;
                        planeOf[k] = p;
                        boxIndexOf[k] = idx;
                    }
        }
        // bounceBackModels3D.hh:48-65 (fluidDirections of a MomentumExchangeBounceBack cell).
        // A neighbour outside a non-periodic domain edge holds the background dynamics
        // in Palabos's envelope, so it counts as fluid.
        for (int x = 0; x < d.nx; ++x) for (int y = 0; y < d.ny; ++y) for (int z = 0; z < d.nz; ++z) {
            long k = lin(x, y, z);
            if (tag[k] != TAG_MEBB) continue;
            uint32_t m = 0;
            for (int i = 1; i < 19; ++i) {
                int nx_ = x + C[i][0], ny_ = y + C[i][1], nz_ = z + C[i][2];
                bool inside = wrap(nx_, ny_, nz_);
                if (!inside) { m |= 1u << i; continue; }
                uint8_t pt = tag[lin(nx_, ny_, nz_)];
                if (pt != TAG_MEBB && pt != TAG_BB && pt != TAG_NODYN && pt != TAG_WALL) m |= 1u << i;
            }
            fluidDirs[k] = m;
        }
    }

    long lin(int x, int y, int z) const { return ((long)x * d.ny + y) * d.nz + z; }

    // Applies periodic wrap; returns false if the point is outside a non-periodic edge.
    bool wrap(int& x, int& y, int& z) const {
        if (x < 0 || x >= d.nx) { if (!d.periodic[0]) return false; x = (x + d.nx) % d.nx; }
        if (y < 0 || y >= d.ny) { if (!d.periodic[1]) return false; y = (y + d.ny) % d.ny; }
        if (z < 0 || z >= d.nz) { if (!d.periodic[2]) return false; z = (z + d.nz) % d.nz; }
        return true;
    }

    // The BGK rates: Palabos's, or exactRates in the 16-bit storage modes.
    Rates<T> bgkRates() const {
        if (d.storage >= 2) {
            const Rates<float> r = exactRates((float)d.omega);
            return {(T)r.one_m_omega, (T)r.t0_omega, (T)r.t1_omega, (T)r.t4_omega, (T)r.omega};
        }
        return palabosRates((T)d.omega);
    }

    // Base dynamics collision, with statistics (isoThermalDynamics.hh:318-330 and :1587-1600).
    void baseCollide(T* f, Stats& st, long k) const {
        T omega = (T)d.omega;
        T uSqr;
        T rhoBar;
        T j[3];
        bool forced = d.force[0] != 0 || d.force[1] != 0 || d.force[2] != 0;
        if (d.alpha || alphaT || forced) {
            uSqr = forcedCollide(f, k, rhoBar, st);
        } else if (d.dynamics == 0) {
            get_rhoBar_j(f, rhoBar, j);
            uSqr = bgk_collision(f, rhoBar, j, bgkRates(), invRhoOf(rhoBar));
        } else {
            T pi[6];
            get_rhoBar_j_PiNeq(f, rhoBar, j, pi);
            T invRho = invRhoOf(rhoBar);
            uSqr = rlb_collision(f, rhoBar, invRho, j, pi, omega);
        }
        // latticeStatistics.h:38-43
        st.sumRhoBar += (double)rhoBar;
        st.sumUSqr += (double)uSqr;
        st.maxUSqr = std::max(st.maxUSqr, (double)uSqr);
        st.count += 1;
    }

    // BGK or RBGK with a Guo body force and/or a Brinkman porosity term.
    // BGK + Guo follows GuoExternalForceBGKdynamics::collide
    // (externalForceDynamics.hh:215-232, velocity :59-74) and addGuoForce
    // (externalForceTemplates3D.h:73-150) term by term. Palabos has no
    // RBGK + Guo and no Brinkman term; they are defined here:
    //   RBGK + Guo: rhoBar, j, PiNeq from compute_rhoBar_j_PiNeq, u = j/rho + F/2,
    //               rlb_collision around j_eq = rho u, then addGuoForce.
    //   Brinkman:   the equilibrium velocity is damped, u <- (1 - alpha) u,
    //               before j_eq = rho u is formed (alpha in [0,1], 1 = solid).
    T forcedCollide(T* f, long k, T& rhoBar, Stats& st) const {
        const Rates<T> rates = bgkRates();
        // BGK in the 16-bit modes: omega' in the drag and the Guo term as well
        T omega = d.dynamics == 0 ? rates.omega : (T)d.omega;
        T j[3];
        T pi[6];
        T rhoBarP;
        if (d.dynamics == 0) {
            rhoBar = get_rhoBar(f);
            get_rhoBar_j(f, rhoBarP, j);
        } else {
            get_rhoBar_j_PiNeq(f, rhoBarP, j, pi);
            rhoBar = rhoBarP;
        }
        T invRhoP = invRhoOf(rhoBarP);
        T F[3] = {(T)d.force[0], (T)d.force[1], (T)d.force[2]};
        T u[3];
        for (int c = 0; c < 3; ++c) u[c] = j[c]*invRhoP + F[c]/(T)2;
        // Brinkman damping; the momentum it removes in the collision is summed as
        // the porous drag: omega*rho*(u - u_damped) for BGK, rho*(u - u_damped)
        // for RBGK, whose first moment becomes j_eq.
        T damp = (T)1 - (alphaT ? alphaT[k] : (d.alpha ? (T)d.alpha[k] : (T)0));
        T rho = rhoBar + (T)1;
        T jeq[3];
        for (int c = 0; c < 3; ++c) {
            T u0 = u[c];
            u[c] = u0 * damp;
            jeq[c] = rho * u[c];
            st.porous[c] += (double)(d.dynamics == 0 ? omega * (rho * u0 - jeq[c]) : (rho * u0 - jeq[c]));
        }
        T uSqr;
        if (d.dynamics == 0) uSqr = bgk_collision(f, rhoBar, jeq, rates, invRhoOf(rhoBar));
        else uSqr = rlb_collision(f, rhoBar, invRhoOf(rhoBar), jeq, pi, omega);
        if (F[0] != 0 || F[1] != 0 || F[2] != 0) addGuoForce(f, u, F, omega);
        return uSqr;
    }

    // Plane boundary completion (dynamics-based or wrapped variant: same arithmetic).
    void completePlane(T* f, long k) const {
        const OraclePlane& P = d.planes[planeOf[k]];
        const long bi = boxIndexOf[k];
        const int dir = P.dir, ori = P.ori;
        // Index subsets (indexTemplates.h:86-133): onWall c[dir]==0, normal c[dir]==ori,
        // outgoing (missing) = opposite(normal).
        int onWall[19], nOnWall = 0, normal[19], nNormal = 0;
        for (int i = 0; i < 19; ++i) {
            if (C[i][dir] == 0) onWall[nOnWall++] = i;
            if (C[i][dir] == ori) normal[nNormal++] = i;
        }
        // computeRhoBarJ: boundaryDynamics.hh:643-700 (velocity) and :776-822 (pressure)
        T rhoBar;
        T j[3];
        if (P.kind == 0) {
            const std::vector<double>& v = vals[planeOf[k]];
            T u[3] = {planeScale * (T)v[3*bi], planeScale * (T)v[3*bi+1], planeScale * (T)v[3*bi+2]};
            T rhoOnWall = T();
            for (int n = 0; n < nOnWall; ++n) rhoOnWall += f[onWall[n]];
            T rhoNormal = T();
            for (int n = 0; n < nNormal; ++n) rhoNormal += f[normal[n]];
            T velNormal = (T)ori * u[dir];
            rhoBar = ((T)2*rhoNormal + rhoOnWall - velNormal) / ((T)1 + velNormal);
            T rho = rhoBar + (T)1;
            for (int c = 0; c < 3; ++c) j[c] = rho * u[c];
        } else {
            rhoBar = (T)vals[planeOf[k]][bi] - (T)1;  // StoreDensityDynamics::defineDensity, boundaryDynamics.hh:251-255
            j[0] = j[1] = j[2] = T();
            T rhoOnWall = T();
            for (int n = 0; n < nOnWall; ++n) rhoOnWall += f[onWall[n]];
            T rhoNormal = T();
            for (int n = 0; n < nNormal; ++n) rhoNormal += f[normal[n]];
            j[dir] = (T)ori * ((T)2*rhoNormal + rhoOnWall - rhoBar);
        }
        T jSqr = j[0]*j[0] + j[1]*j[1] + j[2]*j[2];
        T invRho = invRhoOf(rhoBar);
        if (P.family == 0) {
            // zouHeDynamics.hh:44-107 (ZouHeClosure)
            int missing[5], nMissing = 0;
            for (int n = 0; n < nNormal; ++n) missing[nMissing++] = opposite(normal[n]);
            int diag[5], nDiag = 0;
            bool removed = false;
            for (int n = 0; n < nMissing; ++n) {
                int p = missing[n];
                int nz_ = std::abs(C[p][0]) + std::abs(C[p][1]) + std::abs(C[p][2]);
                if (nz_ == 1 && !removed) { removed = true; continue; }
                diag[nDiag++] = p;
            }
            for (int n = 0; n < nMissing; ++n) {
                int p = missing[n], o = opposite(p);
                f[p] = f[o] - equilibrium(o, rhoBar, invRho, j, jSqr) + equilibrium(p, rhoBar, invRho, j, jSqr);
            }
            // compute_rho_uLb (momentTemplates3D.h:441-449)
            T fRhoBar, fJ[3];
            get_rhoBar_j(f, fRhoBar, fJ);
            T fInvRho = invRhoOf(fRhoBar);
            T falseRho = fRhoBar + (T)1;
            T falseU[3] = {fJ[0]*fInvRho, fJ[1]*fInvRho, fJ[2]*fInvRho};
            T diff[3];
            for (int c = 0; c < 3; ++c) diff[c] = (j[c] - falseRho*falseU[c]) / (T)nDiag;
            for (int n = 0; n < nDiag; ++n)
                for (int iDim = 1; iDim < 3; ++iDim) {
                    int a = (dir + iDim) % 3;
                    f[diag[n]] += C[diag[n]][a] * diff[a];
                }
        } else {
            // regularizedBoundaryDynamics.hh:94-106 / :166-178, boundaryTemplates.h:42-90,
            // isoThermalDynamics.hh:55-69 (regularize)
            T fNeq[19];
            for (int n = 0; n < nOnWall; ++n) {
                int i = onWall[n];
                fNeq[i] = f[i] - equilibrium(i, rhoBar, invRho, j, jSqr);
            }
            for (int n = 0; n < nNormal; ++n) {
                int i = normal[n];
                fNeq[i] = (i == 0) ? T() : f[i] - equilibrium(i, rhoBar, invRho, j, jSqr);
            }
            T pi[6];
            int iPi = 0;
            for (int a = 0; a < 3; ++a)
                for (int b = a; b < 3; ++b) {
                    pi[iPi] = T();
                    for (int n = 0; n < nOnWall; ++n) {
                        int i = onWall[n];
                        pi[iPi] += C[i][a]*C[i][b]*fNeq[i];
                    }
                    for (int n = 0; n < nNormal; ++n) {
                        int i = normal[n];
                        pi[iPi] += (T)2 * C[i][a]*C[i][b]*fNeq[i];
                    }
                    ++iPi;
                }
            f[0] = equilibrium(0, rhoBar, invRho, j, jSqr) + fromPiToFneq(0, pi);
            for (int i = 1; i <= 9; ++i) {
                f[i] = equilibrium(i, rhoBar, invRho, j, jSqr);
                f[i+9] = equilibrium(i+9, rhoBar, invRho, j, jSqr);
                T fn = fromPiToFneq(i, pi);
                f[i] += fn;
                f[i+9] += fn;
            }
        }
    }

    // True if population i of cell (x,y,z) streams from a source that reflects it:
    // outside a non-periodic edge, or a link-wise wall cell.
    bool sourceReflects(int x, int y, int z, int i, long& src) const {
        int sx = x - C[i][0], sy = y - C[i][1], sz = z - C[i][2];
        if (!wrap(sx, sy, sz)) return true;
        src = lin(sx, sy, sz);
        return tag[src] == TAG_WALL;
    }

    // computeRhoBarJ of cell k (BoxRhoBarJfunctional3D, dataAnalysisFunctional3D.hh:528-561):
    // the bulk dynamics' get_rhoBar_j (momentTemplates3D.h:411-423), BounceBack
    // rhoBar(rho) and 0, NoDynamics 1 - SkordosFactor = 0 and 0 (core/dynamics.hh:880-885,
    // :1061-1066). Returns false for dynamics the external processors do not cover here.
    bool rhoBarJ(const T* f, long k, T bbRho, T& rhoBar, T j[3]) const {
        j[0] = j[1] = j[2] = T();
        switch (tag[k]) {
        case TAG_BULK: get_rhoBar_j(f, rhoBar, j); return true;
        case TAG_BB: rhoBar = bbRho - (T)1; return true;
        case TAG_NODYN: rhoBar = T(); return true;
        default: return false;
        }
    }

    // collideExternal of cell k (ExternalRhoJcollideAndStream3D,
    // dynamicsProcessor3D.hh:41-56): BGK with the given moments
    // (isoThermalDynamics.hh:332-338), regularized BGK with PiNeq about them
    // (:1604-1613), BounceBack its collide (core/dynamics.hh:793-798), NoDynamics
    // nothing (:974-977). No statistics.
    bool collideExternal(T* f, long k, T rhoBar, const T j[3]) const {
        switch (tag[k]) {
        case TAG_BULK:
            if (d.alpha || alphaT || d.force[0] != 0 || d.force[1] != 0 || d.force[2] != 0) return false;
            if (d.dynamics == 0) {
                bgk_collision(f, rhoBar, j, bgkRates(), invRhoOf(rhoBar));
            } else {
                T pi[6];
                compute_PiNeq_given(f, rhoBar, j, pi);
                rlb_collision(f, rhoBar, invRhoOf(rhoBar), j, pi, (T)d.omega);
            }
            return true;
        case TAG_BB:
            for (int i = 1; i <= 9; ++i) std::swap(f[i], f[i+9]);
            return true;
        case TAG_NODYN:
            return true;
        default:
            return false;
        }
    }

    // One collision of cell k, in place on f (the cell's 19 populations).
    void collideCell(T* f, long k, Stats& st) const {
        switch (tag[k]) {
        case TAG_BULK:
            baseCollide(f, st, k);
            break;
        case TAG_BB:  // dynamics.hh:783-790
            for (int i = 1; i <= 9; ++i) std::swap(f[i], f[i+9]);
            break;
        case TAG_NODYN:  // dynamics.hh:968-971
        case TAG_WALL:   // solid of link-wise half-way bounce-back: not updated
            break;
        case TAG_MEBB: {  // bounceBackModels.hh:98-135
            T mom[3] = {T(), T(), T()};
            uint32_t m = fluidDirs[k];
            for (int i = 1; i < 19; ++i) {
                if (!(m & (1u << i))) continue;
                int o = opposite(i);
                for (int c = 0; c < 3; ++c) mom[c] += (T)2*C[i][c]*f[o];
            }
            for (int c = 0; c < 3; ++c) st.force[c] += (double)(-mom[c]);
            for (int i = 1; i <= 9; ++i) std::swap(f[i], f[i+9]);
            break;
        }
        case TAG_PLANE:  // core/dynamics.hh:470-475 + :709-711 (completion, then base collision)
            completePlane(f, k);
            baseCollide(f, st, k);
            break;
        case TAG_STORE:  // StoreVelocityDynamics: CompositeDynamics::collide, base collision
                         // (core/dynamics.hh:470-475; boundaryDynamics.hh:171-209)
            baseCollide(f, st, k);
            break;
        }
    }

    // ----- data processors after the stream (MultiBlock3D::executeInternalProcessors):
    // the Neumann-type copies (boundaryCondition/neumannCondition3D.hh:116-252) and
    // the finite-difference edge and corner closures
    // (finiteDifferenceBoundaryProcessor3D.hh:130-329), in the order the boundary
    // objects were added, each object's copy before its closure
    // (boundaryInstantiator3D.h:1017-1247).

    bool hasProcessors() const {
        for (int p = 0; p < d.nplanes; ++p)
            if (d.planes[p].shape != 0 || copyKind(p) != 0) return true;
        return false;
    }

    // 0 none, 1 velocity, 2 tangential velocity, 3 normal velocity, 4 density
    int copyKind(int p) const {
        const OraclePlane& P = d.planes[p];
        const int t = P.bc_type;
        if (P.shape == 0 && P.kind == 1) return t == 1 ? 4 : 0;       // addPressureBoundary
        if (t == 1 || t == 4) return 1;                                // neumann, outflow
        if (t == 2) return 2;                                          // freeslip
        if (t == 5) return 3;                                          // normalOutflow
        return 0;
    }

    void outwardNormal(int p, int n[3]) const {
        const OraclePlane& P = d.planes[p];
        for (int a = 0; a < 3; ++a) n[a] = P.shape == 0 ? (a == P.dir ? P.ori : 0) : P.normal[a];
    }

    // The 19 populations of cell k as its dynamics sees them in a processor: a
    // plane cell of a wrapped factory has been completed by its own processor
    // (wrappedLocalBoundaryProcessor3D.hh:41-62) before any later processor runs.
    template<typename S>
    void cellPops(const S* state, long k, T f[19]) const {
        for (int i = 0; i < 19; ++i) f[i] = Store<S, T>::load(state[k*19 + i]);
        if (tag[k] == TAG_PLANE && d.bc_wrapped) completePlane(f, k);
    }

    // Cell::computeDensity of cell k: bulk and StoreVelocity cells fullRho of
    // get_rhoBar (core/dynamics.hh:278-281, boundaryDynamics.hh:82-87); plane
    // cells fullRho of their Dirichlet rhoBar (boundaryDynamics.hh:640-679,
    // StoreDensityDynamics :229-233).
    template<typename S>
    T densityOf(const S* state, long k) const {
        T f[19];
        cellPops(state, k, f);
        if (tag[k] == TAG_PLANE) {
            T rhoBar, u[3];
            planeMacroscopic(f, k, rhoBar, u);
            return rhoBar + (T)1;
        }
        return get_rhoBar(f) + (T)1;
    }

    // Cell::computeVelocity of cell k: bulk j / rho (compute_uLb with the
    // sequential rhoBar, momentTemplates3D.h:138-144); StoreVelocity and velocity
    // plane cells their stored velocity (boundaryDynamics.hh:324-330); pressure
    // planes j / rho of their closure (:776-835).
    template<typename S>
    void velocityOf(const S* state, long k, T u[3]) const {
        if (tag[k] == TAG_STORE || (tag[k] == TAG_PLANE && d.planes[planeOf[k]].kind == 0)) {
            const double* v = &vals[planeOf[k]][3 * boxIndexOf[k]];
            for (int c = 0; c < 3; ++c) u[c] = (T)v[c];
            return;
        }
        T f[19];
        cellPops(state, k, f);
        if (tag[k] == TAG_PLANE) {
            T rhoBar;
            planeMacroscopic(f, k, rhoBar, u);
            return;
        }
        const T rhoBar = get_rhoBar(f);
        T rb, j[3];
        get_rhoBar_j(f, rb, j);
        const T ir = invRhoOf(rhoBar);
        for (int c = 0; c < 3; ++c) u[c] = j[c] * ir;
    }

    // rhoBar and velocity a plane cell reports (Dirichlet closure from its known populations).
    void planeMacroscopic(const T* f, long k, T& rhoBar, T u[3]) const {
        const OraclePlane& P = d.planes[planeOf[k]];
        const long bi = boxIndexOf[k];
        T onWall = T(), normal = T();
        for (int i = 0; i < 19; ++i) if (C[i][P.dir] == 0) onWall += f[i];
        for (int i = 0; i < 19; ++i) if (C[i][P.dir] == P.ori) normal += f[i];
        if (P.kind == 0) {
            const double* v = &vals[planeOf[k]][3 * bi];
            for (int c = 0; c < 3; ++c) u[c] = (T)v[c];
            const T velNormal = (T)P.ori * u[P.dir];
            rhoBar = ((T)2*normal + onWall - velNormal) / ((T)1 + velNormal);
        } else {
            rhoBar = (T)vals[planeOf[k]][bi] - (T)1;
            T j[3] = {T(), T(), T()};
            j[P.dir] = (T)P.ori * ((T)2*normal + onWall - rhoBar);
            const T ir = invRhoOf(rhoBar);
            for (int c = 0; c < 3; ++c) u[c] = j[c] * ir;
        }
    }

    long neighbour(int x, int y, int z, int dx, int dy, int dz) const {
        int a = x + dx, b = y + dy, c = z + dz;
        if (!wrap(a, b, c)) return -1;
        return lin(a, b, c);
    }

    // IsoThermalBulkDynamics::regularize (isoThermalDynamics.hh:55-69).
    static void regularize(T* f, T rhoBar, const T j[3], T jSqr, const T pi[6]) {
        const T invRho = invRhoOf(rhoBar);
        f[0] = equilibrium(0, rhoBar, invRho, j, jSqr) + fromPiToFneq(0, pi);
        for (int i = 1; i <= 9; ++i) {
            f[i] = equilibrium(i, rhoBar, invRho, j, jSqr);
            f[i+9] = equilibrium(i+9, rhoBar, invRho, j, jSqr);
            T fNeq = fromPiToFneq(i, pi);
            f[i] += fNeq;
            f[i+9] += fNeq;
        }
    }

    // pi from the velocity gradient dA_uB[a][b] = d_a u_b (finiteDifferenceBoundaryProcessor3D.hh:162-172).
    void gradientToPi(const T g[3][3], T rho, T pi[6]) const {
        const T omega = (T)d.omega;
        const T sToPi = -rho / (T)3 / omega;
        pi[0] = (T)2 * g[0][0] * sToPi;
        pi[3] = (T)2 * g[1][1] * sToPi;
        pi[5] = (T)2 * g[2][2] * sToPi;
        pi[1] = (g[0][1] + g[1][0]) * sToPi;
        pi[2] = (g[0][2] + g[2][0]) * sToPi;
        pi[4] = (g[1][2] + g[2][1]) * sToPi;
    }

    template<typename S>
    void storePops(S* state, long k, const T f[19]) const {
        for (int i = 0; i < 19; ++i) state[k*19 + i] = Store<S, T>::store(f[i]);
    }

    // OuterVelocityEdgeFunctional3D::processCell (finiteDifferenceBoundaryProcessor3D.hh:144-195).
    template<typename S>
    bool edgeCell(S* state, int p, int x, int y, int z) const {
        const OraclePlane& P = d.planes[p];
        const int plane = P.dir, d1 = (plane + 1) % 3, d2 = (plane + 2) % 3;
        const int n1 = P.normal[d1], n2 = P.normal[d2];
        int s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0}, sp[3] = {0, 0, 0};
        s1[d1] = -n1; s2[d2] = -n2; sp[plane] = 1;
        const long k = lin(x, y, z);
        const long k10 = neighbour(x, y, z, s1[0], s1[1], s1[2]);
        const long k01 = neighbour(x, y, z, s2[0], s2[1], s2[2]);
        const long kp = neighbour(x, y, z, sp[0], sp[1], sp[2]);
        const long km = neighbour(x, y, z, -sp[0], -sp[1], -sp[2]);
        if (k10 < 0 || k01 < 0 || kp < 0 || km < 0) return false;
        const T rho10 = densityOf(state, k10);
        const T rho01 = densityOf(state, k01);
        const T rho = (T)0.5 * (rho01 + rho10);
        T u0[3], u10[3], u01[3], up[3], um[3];
        velocityOf(state, k, u0);
        velocityOf(state, k10, u10);
        velocityOf(state, k01, u01);
        velocityOf(state, kp, up);
        velocityOf(state, km, um);
        T g[3][3];
        for (int b = 0; b < 3; ++b) {
            g[plane][b] = (up[b] - um[b]) / (T)2;           // fd::ctl_diff
            g[d1][b] = (T)(-n1) * (-u0[b] + u10[b]);        // -orientation * fd::o1_fwd_diff
            g[d2][b] = (T)(-n2) * (-u0[b] + u01[b]);
        }
        T pi[6];
        gradientToPi(g, rho, pi);
        T j[3];
        for (int c = 0; c < 3; ++c) j[c] = rho * u0[c];
        const T jSqr = j[0]*j[0] + j[1]*j[1] + j[2]*j[2];
        T f[19];
        regularize(f, rho - (T)1, j, jSqr, pi);
        storePops(state, k, f);
        return true;
    }

    // OuterVelocityCornerFunctional3D::processCell (finiteDifferenceBoundaryProcessor3D.hh:262-313).
    template<typename S>
    bool cornerCell(S* state, int p, int x, int y, int z) const {
        const int* n = d.planes[p].normal;
        const long k = lin(x, y, z);
        const long k100 = neighbour(x, y, z, -n[0], 0, 0);
        const long k010 = neighbour(x, y, z, 0, -n[1], 0);
        const long k001 = neighbour(x, y, z, 0, 0, -n[2]);
        if (k100 < 0 || k010 < 0 || k001 < 0) return false;
        const T rho100 = densityOf(state, k100);
        const T rho010 = densityOf(state, k010);
        const T rho001 = densityOf(state, k001);
        const T rho = (T)1 / (T)3 * (rho001 + rho010 + rho100);
        T u0[3], ux[3], uy[3], uz[3];
        velocityOf(state, k, u0);
        velocityOf(state, k100, ux);
        velocityOf(state, k010, uy);
        velocityOf(state, k001, uz);
        T g[3][3];
        for (int b = 0; b < 3; ++b) {
            g[0][b] = (T)(-n[0]) * (-u0[b] + ux[b]);
            g[1][b] = (T)(-n[1]) * (-u0[b] + uy[b]);
            g[2][b] = (T)(-n[2]) * (-u0[b] + uz[b]);
        }
        T pi[6];
        gradientToPi(g, rho, pi);
        T j[3];
        for (int c = 0; c < 3; ++c) j[c] = rho * u0[c];
        const T jSqr = j[0]*j[0] + j[1]*j[1] + j[2]*j[2];
        T f[19];
        regularize(f, rho - (T)1, j, jSqr, pi);
        storePops(state, k, f);
        return true;
    }

    // CopyVelocity / CopyTangentialVelocity / CopyNormalVelocity / CopyDensity of one cell.
    template<typename S>
    bool copyCell(const S* state, int p, int kind, int x, int y, int z, long bi) const {
        int n[3];
        outwardNormal(p, n);
        const long src = neighbour(x, y, z, -n[0], -n[1], -n[2]);
        if (src < 0) return false;
        if (kind == 4) {
            vals[p][bi] = (double)densityOf(state, src);
            return true;
        }
        T u[3];
        velocityOf(state, src, u);
        for (int a = 0; a < 3; ++a) {
            if (kind == 2 && n[a] != 0) u[a] = T();
            if (kind == 3 && n[a] == 0) u[a] = T();
        }
        for (int a = 0; a < 3; ++a) vals[p][3 * bi + a] = (double)u[a];
        return true;
    }

    // One pass of every processor; false if a processor's stencil leaves the lattice.
    template<typename S>
    bool runProcessors(S* state) const {
        for (int p = 0; p < d.nplanes; ++p) {
            const OraclePlane& P = d.planes[p];
            const int* b = P.box;
            const int ck = copyKind(p);
            if (ck) {
                long bi = 0;
                for (int x = b[0]; x <= b[1]; ++x) for (int y = b[2]; y <= b[3]; ++y)
                    for (int z = b[4]; z <= b[5]; ++z, ++bi)
                        if (!copyCell(state, p, ck, x, y, z, bi)) return false;
            }
            if (P.shape == 0) continue;
            for (int x = b[0]; x <= b[1]; ++x) for (int y = b[2]; y <= b[3]; ++y)
                for (int z = b[4]; z <= b[5]; ++z) {
                    const bool ok = P.shape == 1 ? edgeCell(state, p, x, y, z) : cornerCell(state, p, x, y, z);
                    if (!ok) return false;
                }
        }
        return true;
    }
};

template<> struct Store<double, double> {
    static double load(double s) { return s; }
    static double store(double v) { return v; }
};
template<> struct Store<float, float> {
    static float load(float s) { return s; }
    static float store(float v) { return v; }
};
struct Half { uint16_t bits; };
struct BHalf { uint16_t bits; };
template<> struct Store<Half, float> {
    static float load(Half s) { return f16_to_f32(s.bits); }
    static Half store(float v) { Half h; h.bits = f32_to_f16(v); return h; }
};
template<> struct Store<BHalf, float> {
    static float load(BHalf s) { return bf16_to_f32(s.bits); }
    static BHalf store(float v) { BHalf h; h.bits = f32_to_bf16(v); return h; }
};

// Streaming of the post-collision populations `post` into `state`, with
// same-step reflection at non-periodic edges (blockLattice3D.hh:403-427,
// latticeTemplates3D.h:50-63, periodicity :554-624). Link-wise walls reflect the
// same way; the force on them is the momentum 2 c_i f*_i of every reflected
// population.
template<typename S, typename T>
void streamAll(const Engine<T>& eng, const std::vector<S>& post, S* state, Stats& st) {
    const OracleDesc& d = eng.d;
    for (int x = 0; x < d.nx; ++x) for (int y = 0; y < d.ny; ++y) for (int z = 0; z < d.nz; ++z) {
        long k = eng.lin(x, y, z);
        if (eng.tag[k] == TAG_WALL) continue;
        state[k*19] = post[k*19];
        for (int i = 1; i < 19; ++i) {
            long src = -1;
            if (!eng.sourceReflects(x, y, z, i, src)) {
                state[k*19 + i] = post[src*19 + i];
            } else {
                state[k*19 + i] = post[k*19 + opposite(i)];
                if (src >= 0) {  // reflected by a wall cell: momentum 2 c_o f*_o, o = opposite(i)
                    int o = opposite(i);
                    T fo = Store<S, T>::load(post[k*19 + o]);
                    for (int c = 0; c < 3; ++c) st.force[c] += (double)((T)2 * C[o][c] * fo);
                }
            }
        }
    }
}

template<typename S, typename T>
int runSteps(const OracleDesc& d, S* state, int nsteps, double* statsOut, bool initialize, bool snapshot = false) {
    Engine<T> eng(d);
    const bool processors = eng.hasProcessors();
    // MultiBlock3D::initialize (multiBlock3D.cpp:330-332): one pass of the processors
    if (initialize && processors && !eng.runProcessors(state)) return -3;
    const long N = eng.N;
    std::vector<S> post((size_t)N * 19);
    Stats st;
    for (int step = 0; step < nsteps; ++step) {
        st = Stats();
        // Collision (every cell reads only its own populations).
        for (long k = 0; k < N; ++k) {
            T f[19];
            for (int i = 0; i < 19; ++i) f[i] = Store<S, T>::load(state[k*19 + i]);
            eng.collideCell(f, k, st);
            for (int i = 0; i < 19; ++i) post[k*19 + i] = Store<S, T>::store(f[i]);
        }
        streamAll<S, T>(eng, post, state, st);
        if (processors && !eng.runProcessors(state)) return -3;
    }
    if (statsOut) {
        statsOut[0] = st.count ? st.sumRhoBar / (double)st.count : 0.0;
        statsOut[1] = st.count ? st.sumUSqr / (double)st.count : 0.0;
        statsOut[2] = st.maxUSqr;
        statsOut[3] = st.force[0];
        statsOut[4] = st.force[1];
        statsOut[5] = st.force[2];
        statsOut[6] = (double)st.count;
        statsOut[7] = st.porous[0];
        statsOut[8] = st.porous[1];
        statsOut[9] = st.porous[2];
    }
    if (snapshot) {
        // what a wrapped factory's processor leaves in the plane cells: the
        // completion with the boundary values as they are now
        for (long k = 0; k < N; ++k) {
            if (eng.tag[k] != TAG_PLANE) continue;
            T f[19];
            for (int i = 0; i < 19; ++i) f[i] = Store<S, T>::load(state[k*19 + i]);
            eng.completePlane(f, k);
            for (int i = 0; i < 19; ++i) state[k*19 + i] = Store<S, T>::store(f[i]);
        }
    }
    return 0;
}

// BoxRhoBarJfunctional3D over the whole lattice: rhoBar and j of every cell, in T,
// widened to double. -4 for cells the external processors do not cover here.
template<typename S, typename T>
int rhoBarJAll(const OracleDesc& d, const S* state, double bbRho, double* rhoBarOut, double* jOut) {
    Engine<T> eng(d);
    for (long k = 0; k < eng.N; ++k) {
        T f[19], rb, j[3];
        for (int i = 0; i < 19; ++i) f[i] = Store<S, T>::load(state[k*19 + i]);
        if (!eng.rhoBarJ(f, k, (T)bbRho, rb, j)) return -4;
        rhoBarOut[k] = (double)rb;
        for (int c = 0; c < 3; ++c) jOut[3*k + c] = (double)j[c];
    }
    return 0;
}

// One ExternalRhoJcollideAndStream3D step with the field (rhoBar, j as double,
// rounded to T as the field's type is T). Periodic lattices only: at a
// non-periodic edge the processor streams through the block's envelope layer
// (its extended domain, dynamicsProcessor3D.hh:118-190), whose cells Palabos
// also collides, so populations enter from outside the domain; that is not
// restated here (-4).
template<typename S, typename T>
int externalStep(const OracleDesc& d, S* state, const double* rhoBar, const double* j) {
    Engine<T> eng(d);
    if (eng.hasProcessors() || !d.periodic[0] || !d.periodic[1] || !d.periodic[2]) return -4;
    std::vector<S> post((size_t)eng.N * 19);
    for (long k = 0; k < eng.N; ++k) {
        T f[19];
        for (int i = 0; i < 19; ++i) f[i] = Store<S, T>::load(state[k*19 + i]);
        const T jj[3] = {(T)j[3*k], (T)j[3*k + 1], (T)j[3*k + 2]};
        if (!eng.collideExternal(f, k, (T)rhoBar[k], jj)) return -4;
        for (int i = 0; i < 19; ++i) post[k*19 + i] = Store<S, T>::store(f[i]);
    }
    Stats st;
    streamAll<S, T>(eng, post, state, st);
    return 0;
}

// Analysis fields of every cell (dataAnalysisWrapper3D.h:695-777), in T:
// computeDensity, computeRhoBar, computeVelocity x y z, computeKineticEnergy,
// computeVelocityNorm. Bulk cells: rho = fullRho(get_rhoBar) (sequential sum,
// momentTemplates3D.h:375-382, core/dynamics.hh:278-281), u = get_j * invRho
// (compute_uLb, momentTemplates3D.h:138-144), or with a Guo force
// j * invRho(rhoBar of get_rhoBar_j) + F/2 (externalForceDynamics.hh:59-74);
// BounceBack / NoDynamics / MomentumExchangeBounceBack: rho = their density
// parameter, rhoBar = rho - 1 (NoDynamics: 0), u = 0 (core/dynamics.hh:814-880,
// 994-1060; bounceBackModels.hh:160-223); link-wise wall solids rho 1, u 0;
// energy normSqr(u) / 2 and norm sqrt(normSqr(u))
// (dataAnalysisFunctional3D.hh:668-719). Returns false for a lattice with
// planes (their closures are not restated here).
template<typename S, typename T>
bool fieldsAll(const OracleDesc& d, const S* state, const double* solidRho, double* out) {
    if (d.nplanes > 0) return false;
    Engine<T> eng(d);
    const bool forced = d.force[0] != 0 || d.force[1] != 0 || d.force[2] != 0;
    for (long k = 0; k < eng.N; ++k) {
        T rho = (T)1, rhoBar = T(), u[3] = {T(), T(), T()};
        const int tag = eng.tag[k];
        if (tag == TAG_BB) { rho = (T)solidRho[0]; rhoBar = rho - (T)1; }
        else if (tag == TAG_NODYN) { rho = (T)solidRho[1]; }
        else if (tag == TAG_MEBB) { rho = (T)solidRho[2]; rhoBar = rho - (T)1; }
        else if (tag == TAG_BULK) {
            T f[19];
            for (int i = 0; i < 19; ++i) f[i] = Store<S, T>::load(state[k*19 + i]);
            rhoBar = get_rhoBar(f);
            T rb, j[3];
            get_rhoBar_j(f, rb, j);
            const T ir = invRhoOf(forced ? rb : rhoBar);
            for (int c = 0; c < 3; ++c) u[c] = forced ? j[c] * ir + (T)d.force[c] / (T)2 : j[c] * ir;
            rho = rhoBar + (T)1;
        }
        const T uSqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        double* o = out + 7 * k;
        o[0] = (double)rho;
        o[1] = (double)rhoBar;
        o[2] = (double)u[0];
        o[3] = (double)u[1];
        o[4] = (double)u[2];
        o[5] = (double)(uSqr / (T)2);
        o[6] = (double)std::sqrt(uSqr);
    }
    return true;
}

template<typename S, typename T>
void completeAll(const OracleDesc& d, S* state) {
    Engine<T> eng(d);
    for (long k = 0; k < eng.N; ++k) {
        if (eng.tag[k] != TAG_PLANE) continue;
        T f[19];
        for (int i = 0; i < 19; ++i) f[i] = Store<S, T>::load(state[k*19 + i]);
        eng.completePlane(f, k);
        for (int i = 0; i < 19; ++i) state[k*19 + i] = Store<S, T>::store(f[i]);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Forward-mode derivatives for the adjoint check: a dual number with M
// derivative components, run through the same Engine (the oracle's own
// collision and streaming), gives exact directional derivatives of the
// time-averaged objective with respect to seeded porosity cells and the
// velocity-plane scale. Independent of the hand-derived adjoint.

template<int M>
struct DualN {
    double v;
    double d[M];
    DualN() : v(0) { for (int k = 0; k < M; ++k) d[k] = 0; }
    DualN(double x) : v(x) { for (int k = 0; k < M; ++k) d[k] = 0; }
    explicit operator double() const { return v; }
};
template<int M> DualN<M> operator+(const DualN<M>& a, const DualN<M>& b) {
    DualN<M> r(a.v + b.v); for (int k = 0; k < M; ++k) r.d[k] = a.d[k] + b.d[k]; return r;
}
template<int M> DualN<M> operator-(const DualN<M>& a, const DualN<M>& b) {
    DualN<M> r(a.v - b.v); for (int k = 0; k < M; ++k) r.d[k] = a.d[k] - b.d[k]; return r;
}
template<int M> DualN<M> operator-(const DualN<M>& a) {
    DualN<M> r(-a.v); for (int k = 0; k < M; ++k) r.d[k] = -a.d[k]; return r;
}
template<int M> DualN<M> operator*(const DualN<M>& a, const DualN<M>& b) {
    DualN<M> r(a.v * b.v); for (int k = 0; k < M; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r;
}
template<int M> DualN<M> operator/(const DualN<M>& a, const DualN<M>& b) {
    DualN<M> r(a.v / b.v);
    for (int k = 0; k < M; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
template<int M> DualN<M> operator*(int a, const DualN<M>& b) { return DualN<M>((double)a) * b; }
template<int M> DualN<M> operator*(const DualN<M>& a, int b) { return a * DualN<M>((double)b); }
template<int M> DualN<M> operator-(int a, const DualN<M>& b) { return DualN<M>((double)a) - b; }
template<int M> DualN<M>& operator+=(DualN<M>& a, const DualN<M>& b) { a = a + b; return a; }
template<int M> DualN<M>& operator-=(DualN<M>& a, const DualN<M>& b) { a = a - b; return a; }
template<int M> DualN<M>& operator*=(DualN<M>& a, const DualN<M>& b) { a = a * b; return a; }
template<int M> bool operator!=(const DualN<M>& a, int b) { return a.v != (double)b; }

constexpr int kSeeds = 16;           // porosity cells carrying a derivative
typedef DualN<kSeeds + 1> DualO;     // + the velocity-plane scale

// j(f) of the adjoint objective on the natural-layout state (see the C ABI).
template<typename T>
T objectiveTerm(const OracleDesc& d, const std::vector<T>& st, const Engine<T>& eng, int xin, int xout,
                double wdp, double wdrag, double wflux) {
    const long plane = (long)d.ny * d.nz;
    T pin = T(), pout = T(), drag = T(), flux = T();
    const T omega = (T)d.omega;
    for (int x = 0; x < d.nx; ++x) for (int y = 0; y < d.ny; ++y) for (int z = 0; z < d.nz; ++z) {
        const long k = eng.lin(x, y, z);
        if (eng.tag[k] == TAG_WALL) continue;   // solid of link-wise walls: not fluid, counts zero
        const T* f = &st[k * 19];
        T rho = (T)1;
        for (int i = 0; i < 19; ++i) rho = rho + f[i];
        T jx = T();
        for (int i = 0; i < 19; ++i) jx = jx + C[i][0] * f[i];
        if (x == xin) pin = pin + rho;
        if (x == xout) pout = pout + rho;
        if (eng.alphaT && eng.tag[k] != TAG_PLANE) drag = drag + omega * eng.alphaT[k] * jx;
        flux = flux + jx;
    }
    const T cs2 = (T)(1.0 / 3.0);
    const double cells = (double)d.nx * d.ny * d.nz;
    return (T)wdp * cs2 * (pin / (T)(double)plane - pout / (T)(double)plane) + (T)wdrag * drag +
           (T)(wflux / cells) * flux;
}

extern "C" {

// Advances `state` (natural layout, AoS [x][y][z][19] f-bar in the storage type)
// by nsteps collide-and-stream steps. statsOut (nullable, 10 doubles) receives the
// last step's average rhoBar, average uSqr, max uSqr (pre-collision moments of
// bulk and plane cells), the force on the obstacle (momentum exchange of
// MomentumExchangeBounceBack nodes plus link-wise walls; x,y,z), the number of
// cells that took statistics and the momentum removed by the Brinkman term
// (x,y,z). Populations of link-wise wall cells are left unchanged.
// Returns 0, or -1 on a bad descriptor.
int oracle_run(const OracleDesc* d, void* state, int nsteps, double* statsOut) {
    if (!d || !state || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || nsteps < 0) return -1;
    switch (d->storage) {
    case 0: return runSteps<double, double>(*d, (double*)state, nsteps, statsOut, false);
    case 1: return runSteps<float, float>(*d, (float*)state, nsteps, statsOut, false);
    case 2: return runSteps<Half, float>(*d, (Half*)state, nsteps, statsOut, false);
    case 3: return runSteps<BHalf, float>(*d, (BHalf*)state, nsteps, statsOut, false);
    }
    return -1;
}

// As oracle_run, with flags: bit 0, one pass of the data processors on the
// initial state first (MultiBlock3D::initialize, multiBlock3D.cpp:330-332);
// bit 1, the plane cells completed at the end with the boundary values of that
// moment (the snapshot a wrapped factory leaves). Returns -3 if a processor's
// stencil leaves the lattice.
int oracle_run_ex(const OracleDesc* d, void* state, int nsteps, double* statsOut, int flags) {
    if (!d || !state || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || nsteps < 0) return -1;
    const bool ini = (flags & 1) != 0, snap = (flags & 2) != 0;
    switch (d->storage) {
    case 0: return runSteps<double, double>(*d, (double*)state, nsteps, statsOut, ini, snap);
    case 1: return runSteps<float, float>(*d, (float*)state, nsteps, statsOut, ini, snap);
    case 2: return runSteps<Half, float>(*d, (Half*)state, nsteps, statsOut, ini, snap);
    case 3: return runSteps<BHalf, float>(*d, (BHalf*)state, nsteps, statsOut, ini, snap);
    }
    return -1;
}

// Applies the plane-boundary completion to the plane cells of `state` in place
// (what Palabos's wrapped boundary processor leaves in the lattice after a step,
// wrappedLocalBoundaryProcessor3D.hh:41-62).
int oracle_complete(const OracleDesc* d, void* state) {
    if (!d || !state) return -1;
    switch (d->storage) {
    case 0: completeAll<double, double>(*d, (double*)state); return 0;
    case 1: completeAll<float, float>(*d, (float*)state); return 0;
    case 2: completeAll<Half, float>(*d, (Half*)state); return 0;
    case 3: completeAll<BHalf, float>(*d, (BHalf*)state); return 0;
    }
    return -1;
}

// Analysis fields of `state` (see fieldsAll): out[cell][7]. solid_rho: density
// parameters of BounceBack, NoDynamics, MomentumExchangeBounceBack. Returns 0,
// -1 on a bad descriptor, -2 for a lattice with planes.
int oracle_fields(const OracleDesc* d, const void* state, const double* solid_rho, double* out) {
    if (!d || !state || !solid_rho || !out) return -1;
    bool ok = false;
    switch (d->storage) {
    case 0: ok = fieldsAll<double, double>(*d, (const double*)state, solid_rho, out); break;
    case 1: ok = fieldsAll<float, float>(*d, (const float*)state, solid_rho, out); break;
    case 2: ok = fieldsAll<Half, float>(*d, (const Half*)state, solid_rho, out); break;
    case 3: ok = fieldsAll<BHalf, float>(*d, (const BHalf*)state, solid_rho, out); break;
    default: return -1;
    }
    return ok ? 0 : -2;
}

// Bulk equilibrium populations f-bar for (rho, u) per cell, FP64
// (dynamicsTemplates3D.h:1994-1997 with j = rho u, cell.hh:124-132).
void oracle_equilibrium(long n, const double* rho, const double* u, double* fOut) {
    for (long k = 0; k < n; ++k) {
        double rhoBar = rho[k] - 1.0;
        double invRho = invRhoOf(rhoBar);
        double j[3] = {rho[k]*u[3*k], rho[k]*u[3*k+1], rho[k]*u[3*k+2]};
        double jSqr = j[0]*j[0] + j[1]*j[1] + j[2]*j[2];
        for (int i = 0; i < 19; ++i) fOut[k*19 + i] = equilibrium(i, rhoBar, invRho, j, jSqr);
    }
}

// J = 1/(T - t0) sum_{t=t0+1..T} j(f^t) of the adjoint objective (pressure drop,
// Brinkman drag, mean x momentum) from FP64 state f0, and its exact derivatives
// with respect to the porosity of 16 seeded cells and to the scale of the
// velocity planes' imposed velocity: out = [J, dJ/dalpha[16], dJ/ds].
int oracle_objective_dual(const OracleDesc* d, const double* f0, int nsteps, int xin, int xout, double wdp,
                          double wdrag, double wflux, int firstStep, const long* seeds, double* out) {
    if (!d || !f0 || nsteps <= 0 || firstStep < 0 || firstStep >= nsteps || d->storage != 0 || !d->alpha) return -1;
    Engine<DualO> eng(*d);
    const long N = eng.N;
    std::vector<DualO> alpha(N);
    for (long k = 0; k < N; ++k) alpha[k] = DualO(d->alpha[k]);
    for (int s = 0; s < kSeeds; ++s) alpha[seeds[s]].d[s] = 1.0;
    eng.alphaT = alpha.data();
    eng.planeScale = DualO(1.0);
    eng.planeScale.d[kSeeds] = 1.0;
    std::vector<DualO> state(N * 19), post(N * 19);
    for (long k = 0; k < N * 19; ++k) state[k] = DualO(f0[k]);
    DualO J = DualO();
    const DualO invT = DualO(1.0 / (nsteps - firstStep));
    Stats st;
    for (int step = 0; step < nsteps; ++step) {
        for (long k = 0; k < N; ++k) {
            DualO f[19];
            for (int i = 0; i < 19; ++i) f[i] = state[k*19 + i];
            eng.collideCell(f, k, st);
            for (int i = 0; i < 19; ++i) post[k*19 + i] = f[i];
        }
        for (int x = 0; x < d->nx; ++x) for (int y = 0; y < d->ny; ++y) for (int z = 0; z < d->nz; ++z) {
            long k = eng.lin(x, y, z);
            state[k*19] = post[k*19];
            for (int i = 1; i < 19; ++i) {
                long src = -1;
                if (!eng.sourceReflects(x, y, z, i, src)) state[k*19 + i] = post[src*19 + i];
                else state[k*19 + i] = post[k*19 + opposite(i)];
            }
        }
        if (step + 1 > firstStep) J = J + invT * objectiveTerm(*d, state, eng, xin, xout, wdp, wdrag, wflux);
    }
    out[0] = J.v;
    for (int k = 0; k <= kSeeds; ++k) out[1 + k] = J.d[k];
    return 0;
}

// BoxRhoBarJfunctional3D of the natural-layout state: rhoBar [N], j [N][3];
// bb_rho: the density parameter of BounceBack cells.
int oracle_rhobar_j(const OracleDesc* d, const void* state, double bb_rho, double* rho_bar, double* j) {
    if (!d || !state || !rho_bar || !j) return -1;
    switch (d->storage) {
    case 0: return rhoBarJAll<double, double>(*d, (const double*)state, bb_rho, rho_bar, j);
    case 1: return rhoBarJAll<float, float>(*d, (const float*)state, bb_rho, rho_bar, j);
    case 2: return rhoBarJAll<Half, float>(*d, (const Half*)state, bb_rho, rho_bar, j);
    case 3: return rhoBarJAll<BHalf, float>(*d, (const BHalf*)state, bb_rho, rho_bar, j);
    }
    return -1;
}

// One ExternalRhoJcollideAndStream3D step of the natural-layout state with the
// field rho_bar [N], j [N][3].
int oracle_external_step(const OracleDesc* d, void* state, const double* rho_bar, const double* j) {
    if (!d || !state || !rho_bar || !j) return -1;
    switch (d->storage) {
    case 0: return externalStep<double, double>(*d, (double*)state, rho_bar, j);
    case 1: return externalStep<float, float>(*d, (float*)state, rho_bar, j);
    case 2: return externalStep<Half, float>(*d, (Half*)state, rho_bar, j);
    case 3: return externalStep<BHalf, float>(*d, (BHalf*)state, rho_bar, j);
    }
    return -1;
}

// The 16-bit modes' BGK constants for omega: [one_m_omega, t0_omega, t1_omega, t4_omega, omega'].
void oracle_exact_rates(float omega, float* out) {
    const Rates<float> r = exactRates(omega);
    out[0] = r.one_m_omega; out[1] = r.t0_omega; out[2] = r.t1_omega; out[3] = r.t4_omega; out[4] = r.omega;
}
void oracle_f32_to_f16(long n, const float* in, uint16_t* out) { for (long k = 0; k < n; ++k) out[k] = f32_to_f16(in[k]); }
void oracle_f16_to_f32(long n, const uint16_t* in, float* out) { for (long k = 0; k < n; ++k) out[k] = f16_to_f32(in[k]); }
void oracle_f32_to_bf16(long n, const float* in, uint16_t* out) { for (long k = 0; k < n; ++k) out[k] = f32_to_bf16(in[k]); }
void oracle_bf16_to_f32(long n, const uint16_t* in, float* out) { for (long k = 0; k < n; ++k) out[k] = bf16_to_f32(in[k]); }

}  // extern "C"
