// setup.cpp — host-side set-up of one cavity: the stretched grid and the per-axis FDM eigensystems.
// Runs once per solver (cached per axis for the eigensystems), in FP64, compiled without contraction so
// the tables equal the host oracle's bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "cavity.h"

namespace cav {

namespace {

// Evenly spaced count values from a to b, last one pinned to b (backend.cu:643-649).
std::vector<double> evenly(double a, double b, int count) {
    std::vector<double> out(count);
    const double h = (b - a) / (count - 1);
    for (int i = 0; i < count; i++) out[i] = i * h + a;
    out[count - 1] = b;
    return out;
}

// Two-sided tanh clustering with first cell dx_wall, beta by 200-step bisection (backend.cu:651-666).
std::vector<double> clustered_faces(int n, double length, double dx_wall) {
    if (dx_wall >= length / n * (1 - 1e-12)) return evenly(0.0, length, n + 1);
    const std::vector<double> xi = evenly(0.0, 1.0, n + 1);
    std::vector<double> f(n + 1);
    auto place = [&](double beta) {
        const double tb = std::tanh(beta);
        for (int i = 0; i <= n; i++) f[i] = length * 0.5 * (1.0 + std::tanh(beta * (2 * xi[i] - 1)) / tb);
    };
    double lo = 1e-8, hi = 30.0;
    for (int it = 0; it < 200; it++) {
        const double mid = 0.5 * (lo + hi);
        place(mid);
        (f[1] - f[0] < dx_wall ? hi : lo) = mid;
    }
    place(0.5 * (lo + hi));
    f[0] = 0.0;
    f[n] = length;
    f[n / 2] = 0.5 * length;
    return f;
}

double aspect(const std::vector<double>& f) {                          // backend.cu:667-671
    double big = 0, small = 1e300;
    for (size_t i = 1; i < f.size(); i++) {
        const double w = f[i] - f[i - 1];
        big = std::max(big, w);
        small = std::min(small, w);
    }
    return big / small;
}

// Cyclic Jacobi on a symmetric row-major matrix (backend.cu:723-746): the same rotation sequence, so the
// eigenvectors (columns of V) and eigenvalues (diagonal of A) equal the reference's FP64 results.
void jacobi(int n, std::vector<double>& A, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    double total = 0;
    for (double q : A) total += q * q;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                if (i != j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
        if (off <= 1e-28 * total) return;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < n; k++) {
                    double& ap = A[(size_t)k * n + p];
                    double& aq = A[(size_t)k * n + q];
                    const double x = ap, y = aq;
                    ap = cs * x - sn * y;
                    aq = sn * x + cs * y;
                }
                for (int k = 0; k < n; k++) {
                    double& ap = A[(size_t)p * n + k];
                    double& aq = A[(size_t)q * n + k];
                    const double x = ap, y = aq;
                    ap = cs * x - sn * y;
                    aq = sn * x + cs * y;
                }
                for (int k = 0; k < n; k++) {
                    double& vp = V[(size_t)k * n + p];
                    double& vq = V[(size_t)k * n + q];
                    const double x = vp, y = vq;
                    vp = cs * x - sn * y;
                    vq = sn * x + cs * y;
                }
            }
    }
    fail("pressure solver set-up: eigenvalues did not converge");
}

Eigen compute_eigen(const Axis& g) {                                   // backend.cu:753-776
    const int n = g.n;
    std::vector<double> A((size_t)n * n, 0.0), V;
    for (int i = 0; i + 1 < n; i++) {                 // face i+1 couples cells i and i+1, Neumann walls
        const double k = 1.0 / g.spacing[i + 1];
        A[(size_t)i * n + i] -= k;
        A[(size_t)(i + 1) * n + i + 1] -= k;
        A[(size_t)i * n + i + 1] += k;
        A[(size_t)(i + 1) * n + i] += k;
    }
    for (int i = 0; i < n; i++)                       // symmetric form W^-1/2 K W^-1/2
        for (int j = 0; j < n; j++) A[(size_t)i * n + j] /= std::sqrt(g.width[i] * g.width[j]);
    jacobi(n, A, V);
    Eigen e;
    e.n = n;
    e.lambda.resize(n);
    for (int i = 0; i < n; i++) e.lambda[i] = A[(size_t)i * n + i];
    e.zero_mode = (int)(std::max_element(e.lambda.begin(), e.lambda.end()) - e.lambda.begin());
    e.lambda[e.zero_mode] = 0.0;
    e.S.resize((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) e.S[(size_t)i * n + j] = V[(size_t)i * n + j] / std::sqrt(g.width[i]);
    return e;
}

}  // namespace

Axis make_axis(const CavConfig& c, int n, double length, bool round_float) {   // backend.cu:672-699
    std::vector<double> f;
    if (!c.grid_stretch) {
        f = clustered_faces(n, length, 1e300);
    } else {
        double d = std::min(c.grid_dx_wall, length / n);
        auto ratio = [&](double dw) { return aspect(clustered_faces(n, length, dw)); };
        if (c.grid_max_aspect > 0 && ratio(d) > c.grid_max_aspect) {
            double lo = d, hi = length / n;
            for (int it = 0; it < 60; it++) {
                const double mid = 0.5 * (lo + hi);
                (ratio(mid) > c.grid_max_aspect ? lo : hi) = mid;
            }
            d = hi;
        }
        f = clustered_faces(n, length, d);
    }
    Axis a;
    a.n = n;
    a.face = f;
    a.width.resize(n);
    a.centre.resize(n);
    a.spacing.resize(n + 1);
    for (int i = 0; i < n; i++) {
        a.width[i] = f[i + 1] - f[i];
        a.centre[i] = 0.5 * (f[i + 1] + f[i]);
    }
    for (int i = 1; i < n; i++) a.spacing[i] = a.centre[i] - a.centre[i - 1];
    a.spacing[0] = 0.5 * a.width[0];
    a.spacing[n] = 0.5 * a.width[n - 1];
    a.length = f[n] - f[0];
    if (round_float)                        // backend.cu:695-697: the FP32 fields use the float-rounded grid
        for (std::vector<double>* v : {&a.face, &a.centre, &a.width, &a.spacing})
            for (double& q : *v) q = (double)(float)q;
    return a;
}

const Eigen& axis_eigen(const Axis& a) {
    static std::mutex mu;
    static std::map<std::string, Eigen> cache;
    std::string key((const char*)a.width.data(), a.width.size() * sizeof(double));
    key.append((const char*)a.spacing.data(), a.spacing.size() * sizeof(double));
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    Eigen e = compute_eigen(a);
    std::lock_guard<std::mutex> lock(mu);
    return cache.emplace(key, std::move(e)).first->second;
}

// Folded halves of the symmetric form A = W^-1/2 K W^-1/2: A_e/o[i][j] = A[i][j] +/- A[i][n-1-j] (i, j < h).
// Unit eigenvectors v of A_e/o unfold to unit eigenvectors (v, +/- rev v)/sqrt 2 of A with definite parity,
// also inside near-degenerate pairs of wall-localised modes, where a full-matrix Jacobi mixes the two.
SymSplit symmetric_split(const Axis& a) {
    SymSplit sp;
    const int n = a.n;
    if (n % 2 || n < 4) return sp;
    double wmax = 0;
    for (double w : a.width) wmax = std::max(wmax, w);
    for (int i = 0; i < n; i++)
        if (std::fabs(a.width[i] - a.width[n - 1 - i]) > 1e-12 * wmax) return sp;
    for (int i = 0; i <= n; i++)
        if (std::fabs(a.spacing[i] - a.spacing[n - i]) > 1e-12 * wmax) return sp;
    const int h = n / 2;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i + 1 < n; i++) {                 // the 1-D operator of backend.cu:757-762
        const double k = 1.0 / a.spacing[i + 1];
        A[(size_t)i * n + i] -= k;
        A[(size_t)(i + 1) * n + i + 1] -= k;
        A[(size_t)i * n + i + 1] += k;
        A[(size_t)(i + 1) * n + i] += k;
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) A[(size_t)i * n + j] /= std::sqrt(a.width[i] * a.width[j]);
    std::vector<double> lam[2], vec[2];
    for (int par = 0; par < 2; par++) {
        const double sg = par ? -1.0 : 1.0;
        std::vector<double> H((size_t)h * h), V;
        for (int i = 0; i < h; i++)
            for (int j = 0; j < h; j++) H[(size_t)i * h + j] = A[(size_t)i * n + j] + sg * A[(size_t)i * n + (n - 1 - j)];
        jacobi(h, H, V);
        lam[par].resize(h);
        for (int j = 0; j < h; j++) lam[par][j] = H[(size_t)j * h + j];
        vec[par] = V;
    }
    // the constant (zero) mode is the largest eigenvalue overall; it lies in the even half
    int zero = 0;
    for (int j = 1; j < h; j++) if (lam[0][j] > lam[0][zero]) zero = j;
    for (int j = 0; j < h; j++)
        if (lam[1][j] > lam[0][zero]) return sp;      // not the Neumann operator's structure: stay dense
    sp.h = h;
    sp.zero = zero;
    sp.lambda.resize(n);
    for (int j = 0; j < h; j++) { sp.lambda[j] = lam[0][j]; sp.lambda[h + j] = lam[1][j]; }
    sp.lambda[zero] = 0.0;
    sp.fe.assign((size_t)h * h, 0.0);
    sp.fo.assign((size_t)h * h, 0.0);
    sp.be.assign((size_t)h * h, 0.0);
    sp.bo.assign((size_t)h * h, 0.0);
    const double r2 = 1.0 / std::sqrt(2.0);
    for (int j = 0; j < h; j++)
        for (int i = 0; i < h; i++) {
            // S[i][mode] = v[i] / sqrt(w_i) for the unfolded unit eigenvector v
            const double se = vec[0][(size_t)i * h + j] * r2 / std::sqrt(a.width[i]);
            const double so = vec[1][(size_t)i * h + j] * r2 / std::sqrt(a.width[i]);
            sp.fe[(size_t)j * h + i] = se;
            sp.fo[(size_t)j * h + i] = so;
            sp.be[(size_t)i * h + j] = se;
            sp.bo[(size_t)i * h + j] = so;
        }
    sp.ok = true;
    return sp;
}

int pad32(int n) { return (n + 31) / 32 * 32; }

// backend.cu:58-65: the lid profile's edge factor (1 - s^18)^2, in the kernels' association
static double lid_edge(double s) {
    const double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4, s16 = s8 * s8, a = 1.0 - s16 * s2;
    return a * a;
}

std::vector<double> axis_tables(const Axis& a, double K, bool lid_shaped, bool f32) {
    const int n = a.n, L = n + 1;
    std::vector<double> t((size_t)TAB_KINDS * L, 0.0);
    auto at = [&](int kind, int e) -> double& { return t[(size_t)kind * L + e]; };
    if (f32) {
        // The FP32 fields: the reference's float grid quantities and divisors, formed in float as its kernels
        // form them (backend.cu:41-45, 58-65, 108-138, 279-330); the kinds named for a reciprocal hold the
        // divisor itself (the kernels divide in float). The grid is already rounded through float.
        const float len = (float)a.length, Kf = (float)K;
        auto lid_edge_f = [](float s) {
            float s2 = s * s, s4 = s2 * s2, s8 = s4 * s4, s16 = s8 * s8, q = 1.0f - s16 * s2;
            return q * q;
        };
        for (int e = 0; e <= n; e++) {
            at(T_LIDF, e) = lid_shaped ? lid_edge_f(2.0f * ((float)a.face[e] / len) - 1.0f) : 1.0;
            at(T_LIDC, e) = lid_shaped && e < n ? lid_edge_f(2.0f * ((float)a.centre[e] / len) - 1.0f) : 1.0;
        }
        for (int e = 0; e < n; e++) {
            at(T_W, e) = (float)a.width[e];
            at(T_RW, e) = (float)a.width[e];
            at(T_CE, e) = (float)a.centre[e];
        }
        for (int e = 0; e <= n; e++) {
            const float sp = (float)a.spacing[e];
            at(T_RSPD, e) = 1.0 / a.spacing[e];
            at(T_SP, e) = sp;
            at(T_RSP, e) = sp;
            at(T_R2SP, e) = 2.0f * sp;
            const bool wide = e - 2 >= 0 && e + 1 <= n - 1;
            const float ew = wide ? 0.5f * ((float)a.width[e - 2] + (float)a.width[e + 1]) : sp;
            at(T_C3, e) = ew / (2.0f * sp);
            at(T_RGAP3, e) = wide ? (float)a.centre[e + 1] - (float)a.centre[e - 2] : 1.0f;
            at(T_RWS3, e) = wide ? (float)a.width[e - 2] + (float)a.width[e + 1] : 1.0f;
            const float d = (float)(e < n - e ? e : n - e) - 1.0f;
            at(T_RAMP, e) = std::fmin(std::fmax(d / Kf, 0.0f), 1.0f);
        }
        return t;
    }
    for (int e = 0; e <= n; e++) {
        at(T_LIDF, e) = lid_shaped ? lid_edge(2.0 * (a.face[e] / a.length) - 1.0) : 1.0;
        at(T_LIDC, e) = lid_shaped && e < n ? lid_edge(2.0 * (a.centre[e] / a.length) - 1.0) : 1.0;
    }
    for (int e = 0; e < n; e++) {
        at(T_W, e) = a.width[e];
        at(T_RW, e) = 1.0 / a.width[e];
        at(T_CE, e) = a.centre[e];
    }
    for (int e = 0; e <= n; e++) {
        at(T_SP, e) = a.spacing[e];
        at(T_RSP, e) = 1.0 / a.spacing[e];
        at(T_RSPD, e) = at(T_RSP, e);
        at(T_R2SP, e) = 1.0 / (2.0 * a.spacing[e]);
        const bool wide = e - 2 >= 0 && e + 1 <= n - 1;
        at(T_C3, e) = (wide ? 0.5 * (a.width[e - 2] + a.width[e + 1]) : a.spacing[e]) / (2.0 * a.spacing[e]);
        if (wide) {
            at(T_RGAP3, e) = 1.0 / (a.centre[e + 1] - a.centre[e - 2]);
            at(T_RWS3, e) = 1.0 / (a.width[e - 2] + a.width[e + 1]);
        }
        const double d = (double)(e < n - e ? e : n - e) - 1.0;
        at(T_RAMP, e) = std::fmin(std::fmax(d / K, 0.0), 1.0);
    }
    return t;
}

template <class T>
std::vector<T> pack_fragments(const std::vector<double>& M, int n) {
    const int n32 = pad32(n), kts = n32 / 4, jts = n32 / 16;
    std::vector<T> out((size_t)jts * kts * 64, T(0));
    for (int jt = 0; jt < jts; jt++)
        for (int kt = 0; kt < kts; kt++)
            for (int l = 0; l < 64; l++) {
                const int j = 16 * jt + l % 16, i = 4 * kt + l / 16;
                if (j < n && i < n) out[((size_t)jt * kts + kt) * 64 + l] = (T)M[(size_t)j * n + i];
            }
    return out;
}

template std::vector<double> pack_fragments<double>(const std::vector<double>&, int);

namespace {
// float -> bf16 bits, round to nearest even (finite inputs)
uint16_t to_bf16(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
float from_bf16(uint16_t h) {
    uint32_t b = (uint32_t)h << 16;
    float x;
    std::memcpy(&x, &b, 4);
    return x;
}
}  // namespace

void pack_bf16_fragments(const std::vector<double>& M, int n, std::vector<uint16_t>& hi, std::vector<uint16_t>& lo) {
    const int n32 = (n + 31) / 32 * 32, kts = n32 / 32, jts = n32 / 16;
    hi.assign((size_t)jts * kts * 64 * 8, 0);
    lo.assign(hi.size(), 0);
    for (int jt = 0; jt < jts; jt++)
        for (int kt = 0; kt < kts; kt++)
            for (int l = 0; l < 64; l++)
                for (int q = 0; q < 8; q++) {
                    const int j = 16 * jt + l % 16, i = 32 * kt + 8 * (l / 16) + q;
                    if (j >= n || i >= n) continue;
                    const double v = M[(size_t)j * n + i];
                    const uint16_t h = to_bf16((float)v);
                    const size_t at = (((size_t)jt * kts + kt) * 64 + l) * 8 + q;
                    hi[at] = h;
                    lo[at] = to_bf16((float)(v - (double)from_bf16(h)));
                }
}
template std::vector<float> pack_fragments<float>(const std::vector<double>&, int);

std::vector<double> level_records(const std::vector<double>& ztab, int nz) {
    const int L1 = nz + 1;
    auto tv = [&](int kind, int e, int hi) { return ztab[(size_t)kind * L1 + (e < 0 ? 0 : e > hi ? hi : e)]; };
    std::vector<double> r((size_t)(nz + 4) * ZREC, 0.0);
    for (int L = -1; L <= nz + 2; L++) {
        double* q = &r[(size_t)(L + 1) * ZREC];
        q[ZR_RAMP] = tv(T_RAMP, L, nz);
        q[ZR_RSP] = tv(T_RSP, L, nz);
        q[ZR_RGAP3] = tv(T_RGAP3, L, nz);
        q[ZR_R2SP] = tv(T_R2SP, L, nz);
        q[ZR_RWS3] = tv(T_RWS3, L, nz);
        q[ZR_WM2] = tv(T_W, L - 2, nz - 1);
        q[ZR_WM1] = tv(T_W, L - 1, nz - 1);
        q[ZR_W0] = tv(T_W, L, nz - 1);
        q[ZR_WP1] = tv(T_W, L + 1, nz - 1);
        q[ZR_C3] = tv(T_C3, L, nz);
        q[ZR_RW] = tv(T_RW, L, nz - 1);
        q[ZR_RWM1] = tv(T_RW, L - 1, nz - 1);
        const int j = L - 2;
        q[ZR_LIDCJ] = tv(T_LIDC, j, nz);
        q[ZR_RSPT] = tv(T_RSP, j < nz - 1 ? j + 1 : nz, nz);
        q[ZR_RSPJ] = tv(T_RSP, j, nz);
        q[ZR_RWJ] = tv(T_RW, j, nz - 1);
    }
    return r;
}

}  // namespace cav
