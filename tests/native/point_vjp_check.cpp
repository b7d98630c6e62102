// Host-side check of the hand-written reverse-mode point kernels against
// central finite differences (test infrastructure; built and run by
// tests/test_point_vjp.py).  For random states in every direction it
// compares  sum_f mu_f d out_f / d Q_e  from the VJP with the central
// difference of the forward function, for every input component e, and
// prints the largest relative error per case.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../hermeshd_amd/csrc/hd_physics.h"

using namespace hd;

// Roe sound speed squared of the legacy averages (roe_point's branch test)
template <int DIR>
static double roe_asq(const Phys& ph, const double* QL, const double* QR) {
    int ipar, ip1, ip2;
    dir_fields<DIR>(ipar, ip1, ip2);
    const double cp = ph.gamma / ph.gm1;
    double sk[2], sw[2][4];
    for (int s = 0; s < 2; s++) {
        const double* Q = s ? QR : QL;
        const double a = Q[ipar] / Q[RH], b = Q[ip1] / Q[RH], c = Q[ip2] / Q[RH];
        const double vsq = a * a + b * b + c * c, Pri = ph.gm1 * (Q[EN] - 0.5 * Q[RH] * vsq);
        sk[s] = std::sqrt(Q[RH]);
        sw[s][0] = sk[s] * a; sw[s][1] = sk[s] * b; sw[s][2] = sk[s] * c;
        sw[s][3] = 0.5 * sk[s] * (vsq + cp * Pri / Q[RH]);
    }
    const double dsk = 0.5 * (sk[0] + sk[1]);
    double v[3], vcsq = 0.;
    for (int k = 0; k < 3; k++) { v[k] = 0.5 * (sw[0][k] + sw[1][k]) / dsk; vcsq += v[k] * v[k]; }
    return ph.gm1 * (0.5 * (sw[0][3] + sw[1][3]) / dsk - 0.5 * vcsq);
}

// force_llf: the left state has a negative pressure, deep enough that the
// Roe sound speed squared is not positive (the Lax-Friedrichs fallback, whose
// freezing speed comes from the right state); cases that stay on the Roe
// branch are redrawn.
template <int DIR>
static double roe_case(const Phys& ph, std::mt19937_64& gen, bool force_llf) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    double QL[5], QR[5];
    for (int s = 0; s < 2; s++) {
        double* Q = s ? QR : QL;
        const double rho = 0.4 + 1.2 * U(gen);
        const double p = (force_llf && s == 0) ? -(2.0 + 3.0 * U(gen)) : 0.3 + 1.5 * U(gen);
        const double v[3] = {-1.5 + 3 * U(gen), -1.5 + 3 * U(gen), -1.5 + 3 * U(gen)};
        Q[RH] = rho;
        for (int a = 0; a < 3; a++) Q[MX + a] = rho * v[a];
        Q[EN] = p / ph.gm1 + 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    if (force_llf && roe_asq<DIR>(ph, QL, QR) > 0.) return roe_case<DIR>(ph, gen, force_llf);
    if (!force_llf && !(roe_asq<DIR>(ph, QL, QR) > 0.)) return roe_case<DIR>(ph, gen, force_llf);
    double mu[5];
    for (int f = 0; f < 5; f++) mu[f] = -1 + 2 * U(gen);
    const Phys& p2 = ph;
    auto F = [&](const double* L, const double* R, double* out) {
        double FL[5], FR[5];
        flux_point<5, DIR>(ph, L, FL, nullptr);
        flux_point<5, DIR>(ph, R, FR, nullptr);
        for (int f = 0; f < 5; f++) out[f] = 0.5 * (FL[f] + FR[f]);
        roe_point<DIR, 5>(force_llf ? p2 : ph, L, R, FL, FR, out);
        // chain through the physical fluxes as the adjoint does
    };
    double FL[5], FR[5];
    flux_point<5, DIR>(ph, QL, FL, nullptr);
    flux_point<5, DIR>(ph, QR, FR, nullptr);
    double bQL[5] = {0}, bQR[5] = {0}, bFL[5] = {0}, bFR[5] = {0};
    roe_point_vjp<DIR>(force_llf ? p2 : ph, QL, QR, mu, bQL, bQR, bFL, bFR);
    // the physical flux's own dependence: finite differences of flux_point weighted by bF
    double worst = 0.;
    for (int side = 0; side < 2; side++)
        for (int e = 0; e < 5; e++) {
            double L[5], R[5];
            for (int f = 0; f < 5; f++) { L[f] = QL[f]; R[f] = QR[f]; }
            double* X = side ? R : L;
            const double h = 1e-6 * (1.0 + std::fabs(X[e]));
            const double x0 = X[e];
            double op[5], om[5], fp_[5], fm_[5];
            X[e] = x0 + h; F(L, R, op);
            flux_point<5, DIR>(ph, side ? R : L, fp_, nullptr);
            X[e] = x0 - h; F(L, R, om);
            flux_point<5, DIR>(ph, side ? R : L, fm_, nullptr);
            X[e] = x0;
            double fd = 0., fdF = 0.;
            const double* bF = side ? bFR : bFL;
            for (int f = 0; f < 5; f++) {
                fd += mu[f] * (op[f] - om[f]) / (2 * h);
                fdF += bF[f] * (fp_[f] - fm_[f]) / (2 * h);
            }
            // out depends on Q directly (Roe dissipation) and through F(Q)
            const double got = (side ? bQR : bQL)[e] + fdF;
            const double err = std::fabs(got - fd) / (1e-3 + std::fabs(fd));
            if (!(err <= worst)) worst = err;  // a NaN counts as the worst error
        }
    return worst;
}

int main() {
    Phys ph{};
    ph.gamma = 1.4; ph.gm1 = 0.4; ph.P_floor = 1e-12; ph.rh_floor = 1e-12;
    std::mt19937_64 gen(12345);
    double w[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for (int llf = 0; llf < 2; llf++)
        for (int t = 0; t < 300; t++) {
            w[llf][0] = std::fmax(w[llf][0], roe_case<0>(ph, gen, llf));
            w[llf][1] = std::fmax(w[llf][1], roe_case<1>(ph, gen, llf));
            w[llf][2] = std::fmax(w[llf][2], roe_case<2>(ph, gen, llf));
        }
    std::printf("roe %.3e %.3e %.3e\n", w[0][0], w[0][1], w[0][2]);
    std::printf("roe_llf %.3e %.3e %.3e\n", w[1][0], w[1][1], w[1][2]);
    return 0;
}
