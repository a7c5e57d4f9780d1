// Gaussian integrals for the SCF inputs: overlap, kinetic energy, nuclear
// attraction, and the two- and three-centre Coulomb integrals of density
// fitting, over contracted real solid-harmonic Gaussians (McMurchie-
// Davidson: Hermite expansion coefficients E, Hermite Coulomb integrals R
// from the Boys function). Host C++, parallel over shell pairs with
// std::thread; written from the mathematics (Helgaker, Jorgensen and Olsen,
// Molecular Electronic-Structure Theory, ch. 6 and 9).
//
// Conventions (those of the reference, hafnium/basis/shell.py:1-11 and its
// ADR-0006/0009): spherical functions m = -l..l except p = (x, y, z), real
// solid harmonics with the Helgaker-Jorgensen-Olsen phases; functions of a
// shell contraction-major, then m; contraction coefficients multiply
// normalised primitives and every contracted function has unit norm.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "hfjk_internal.h"

namespace {

constexpr int kMaxL = 5;           // up to h functions
constexpr int kMaxBoys = 4 * kMaxL + 4;
constexpr double kPi = 3.14159265358979323846;
constexpr double kPairCut = 1e-22;  // primitive pairs with a smaller Gaussian-product prefactor are skipped

int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian components of angular momentum l: x power descending, then y.
void cart_list(int l, std::vector<int>& ix, std::vector<int>& iy, std::vector<int>& iz) {
  ix.clear();
  iy.clear();
  iz.clear();
  for (int i = l; i >= 0; --i)
    for (int j = l - i; j >= 0; --j) {
      ix.push_back(i);
      iy.push_back(j);
      iz.push_back(l - i - j);
    }
}

double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

double factorial(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}

// Spherical transform: sph[(m index)][cart index] for the 2l+1 functions in
// the library order (p: x, y, z; else m = -l..l).
std::vector<double> cart_to_sph(int l) {
  std::vector<int> ix, iy, iz;
  cart_list(l, ix, iy, iz);
  const int nc = ncart(l);
  std::vector<double> out((2 * l + 1) * nc, 0.0);
  auto index_of = [&](int a, int b, int c) {
    for (int k = 0; k < nc; ++k)
      if (ix[k] == a && iy[k] == b && iz[k] == c) return k;
    throw std::logic_error("cartesian component not found");
  };
  std::vector<int> order;
  if (l == 1) order = {1, -1, 0};
  else
    for (int m = -l; m <= l; ++m) order.push_back(m);
  for (int row = 0; row < (int)order.size(); ++row) {
    const int m = order[row];
    const int am = std::abs(m);
    const int wm = m < 0 ? 1 : 0;
    const double norm = 1.0 / (std::pow(2.0, am) * factorial(l)) *
                        std::sqrt(2.0 * factorial(l + am) * factorial(l - am) / (m == 0 ? 2.0 : 1.0));
    for (int t = 0; t <= (l - am) / 2; ++t)
      for (int u = 0; u <= t; ++u)
        for (int w = wm; w <= am; w += 2) {
          const double sign = ((t + (w - wm) / 2) % 2) ? -1.0 : 1.0;
          const double c = sign * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) * binom(t, u) * binom(am, w);
          const int px = 2 * t + am - 2 * u - w, py = 2 * u + w, pz = l - 2 * t - am;
          out[row * nc + index_of(px, py, pz)] += norm * c;
        }
  }
  return out;
}

const std::vector<double>& sph_table(int l) {
  static std::vector<double> tables[kMaxL + 1];
  static bool ready[kMaxL + 1] = {false};
  if (!ready[l]) {
    tables[l] = cart_to_sph(l);
    ready[l] = true;
  }
  return tables[l];
}

// Boys function F_n(T) for n = 0..nmax.
void boys(int nmax, double t, double* f) {
  if (t < 1e-14) {
    for (int n = 0; n <= nmax; ++n) f[n] = 1.0 / (2 * n + 1) - t / (2 * n + 3);
    return;
  }
  const double et = std::exp(-t);
  if (t < 45.0) {
    // Series for the highest order, then downward recursion.
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * t / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    f[nmax] = et * sum;
    for (int n = nmax - 1; n >= 0; --n) f[n] = (2.0 * t * f[n + 1] + et) / (2 * n + 1);
  } else {
    f[0] = 0.5 * std::sqrt(kPi / t) * std::erf(std::sqrt(t));
    for (int n = 0; n < nmax; ++n) f[n + 1] = ((2 * n + 1) * f[n] - et) / (2.0 * t);
  }
}

// Hermite expansion coefficients in one dimension: e[(i * (lb + 1) + j) * (la + lb + 1) + t].
struct Herm {
  int la, lb, nt;
  std::vector<double> e;
  double& at(int i, int j, int t) { return e[(i * (lb + 1) + j) * nt + t]; }
  double get(int i, int j, int t) const {
    if (t < 0 || t > i + j) return 0.0;
    return e[(i * (lb + 1) + j) * nt + t];
  }
};

void hermite(int la, int lb, double a, double b, double xab, Herm& h) {
  h.la = la;
  h.lb = lb;
  h.nt = la + lb + 1;
  h.e.assign((la + 1) * (lb + 1) * h.nt, 0.0);
  const double p = a + b, mu = a * b / p;
  const double xpa = -b / p * xab, xpb = a / p * xab, i2p = 0.5 / p;
  h.at(0, 0, 0) = std::exp(-mu * xab * xab);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0)
          v = i2p * h.get(i - 1, j, t - 1) + xpa * h.get(i - 1, j, t) + (t + 1) * h.get(i - 1, j, t + 1);
        else
          v = i2p * h.get(i, j - 1, t - 1) + xpb * h.get(i, j - 1, t) + (t + 1) * h.get(i, j - 1, t + 1);
        h.at(i, j, t) = v;
      }
    }
}

// Hermite Coulomb integrals R_tuv = R^0_tuv(p, X, Y, Z) for t + u + v <= lmax.
struct RTab {
  int l;
  std::vector<double> r;  // [(t * (l + 1) + u) * (l + 1) + v]
  double get(int t, int u, int v) const { return r[(t * (l + 1) + u) * (l + 1) + v]; }
};

void hermite_coulomb(int lmax, double p, double x, double y, double z, RTab& out) {
  const int n1 = lmax + 1;
  double f[kMaxBoys + 1];
  boys(lmax, p * (x * x + y * y + z * z), f);
  // work[n][t][u][v]
  std::vector<double> w((size_t)n1 * n1 * n1 * n1, 0.0);
  auto W = [&](int n, int t, int u, int v) -> double& { return w[((n * n1 + t) * n1 + u) * n1 + v]; };
  double fac = 1.0;
  for (int n = 0; n <= lmax; ++n) {
    W(n, 0, 0, 0) = fac * f[n];
    fac *= -2.0 * p;
  }
  for (int n = lmax - 1; n >= 0; --n) {
    const int top = lmax - n;
    for (int t = 0; t <= top; ++t)
      for (int u = 0; u <= top - t; ++u)
        for (int v = 0; v <= top - t - u; ++v) {
          if (t + u + v == 0) continue;
          double val;
          if (t > 0)
            val = (t > 1 ? (t - 1) * W(n + 1, t - 2, u, v) : 0.0) + x * W(n + 1, t - 1, u, v);
          else if (u > 0)
            val = (u > 1 ? (u - 1) * W(n + 1, t, u - 2, v) : 0.0) + y * W(n + 1, t, u - 1, v);
          else
            val = (v > 1 ? (v - 1) * W(n + 1, t, u, v - 2) : 0.0) + z * W(n + 1, t, u, v - 1);
          W(n, t, u, v) = val;
        }
  }
  out.l = lmax;
  out.r.assign((size_t)n1 * n1 * n1, 0.0);
  for (int t = 0; t <= lmax; ++t)
    for (int u = 0; u <= lmax - t; ++u)
      for (int v = 0; v <= lmax - t - u; ++v) out.r[(t * n1 + u) * n1 + v] = W(0, t, u, v);
}

struct Shell {
  int l;
  double c[3];
  std::vector<double> exps;
  std::vector<double> coefs;  // [ncontr][nprim], bare-primitive scale (radial factor included)
  int ncontr;
  int offset;  // first function
  int nfunc() const { return ncontr * (2 * l + 1); }
};

std::vector<Shell> make_shells(const hfjk_basis* b, int& nfunc) {
  if (b == nullptr || b->nshell < 0) throw std::invalid_argument("invalid basis");
  std::vector<Shell> out(b->nshell);
  int eo = 0, co = 0, f = 0;
  for (int s = 0; s < b->nshell; ++s) {
    Shell& sh = out[s];
    sh.l = b->l[s];
    if (sh.l < 0 || sh.l > kMaxL) throw std::invalid_argument("angular momentum beyond the supported l <= 5");
    const int np = b->nprim[s];
    sh.ncontr = b->ncontr[s];
    if (np < 1 || sh.ncontr < 1) throw std::invalid_argument("a shell needs primitives and contractions");
    for (int k = 0; k < 3; ++k) sh.c[k] = b->centers[3 * s + k];
    sh.exps.assign(b->exps + eo, b->exps + eo + np);
    sh.coefs.resize((size_t)sh.ncontr * np);
    for (int ci = 0; ci < sh.ncontr; ++ci)
      for (int k = 0; k < np; ++k) {
        // Coefficients multiply normalised primitives: the radial factor
        // (2a)^((2l+3)/4) carries the exponent dependence of the norm; the
        // l-dependent constant cancels in the final normalisation.
        sh.coefs[(size_t)ci * np + k] = b->coefs[co + ci * np + k] * std::pow(2.0 * sh.exps[k], (2.0 * sh.l + 3.0) / 4.0);
      }
    eo += np;
    co += sh.ncontr * np;
    sh.offset = f;
    f += sh.nfunc();
  }
  nfunc = f;
  return out;
}

// Cartesian primitive-contracted integrals of a shell pair for the
// one-electron operators. kind 0 overlap, 1 kinetic, 2 nuclear attraction.
// out[(ca * ncart_a + ia) * (ncb * ncart_b) + cb * ncart_b + ib] (ca, cb contractions).
void pair_1e(const Shell& A, const Shell& B, int kind, int natom, const double* z, const double* xyz,
             std::vector<double>& out) {
  const int la = A.l, lb = B.l, na = ncart(la), nb = ncart(lb);
  std::vector<int> ax, ay, az, bx, by, bz;
  cart_list(la, ax, ay, az);
  cart_list(lb, bx, by, bz);
  const int npa = (int)A.exps.size(), npb = (int)B.exps.size();
  out.assign((size_t)A.ncontr * na * B.ncontr * nb, 0.0);
  std::vector<double> prim(na * nb);
  const double abx = A.c[0] - B.c[0], aby = A.c[1] - B.c[1], abz = A.c[2] - B.c[2];
  const int extra = kind == 1 ? 2 : 0;
  Herm hx, hy, hz;
  RTab rt;
  for (int i = 0; i < npa; ++i)
    for (int j = 0; j < npb; ++j) {
      const double a = A.exps[i], b = B.exps[j], p = a + b;
      const double mu = a * b / p;
      if (std::exp(-mu * (abx * abx + aby * aby + abz * abz)) < kPairCut) continue;
      hermite(la, lb + extra, a, b, abx, hx);
      hermite(la, lb + extra, a, b, aby, hy);
      hermite(la, lb + extra, a, b, abz, hz);
      const double sq = std::sqrt(kPi / p);
      if (kind == 0) {
        for (int u = 0; u < na; ++u)
          for (int v = 0; v < nb; ++v)
            prim[u * nb + v] = hx.get(ax[u], bx[v], 0) * hy.get(ay[u], by[v], 0) * hz.get(az[u], bz[v], 0) * sq * sq * sq;
      } else if (kind == 1) {
        auto s1 = [&](const Herm& h, int ia, int jb) { return jb < 0 ? 0.0 : h.get(ia, jb, 0) * sq; };
        auto d2 = [&](const Herm& h, int ia, int jb) {
          return jb * (jb - 1) * s1(h, ia, jb - 2) - 2.0 * b * (2 * jb + 1) * s1(h, ia, jb) + 4.0 * b * b * s1(h, ia, jb + 2);
        };
        for (int u = 0; u < na; ++u)
          for (int v = 0; v < nb; ++v) {
            const double sx = s1(hx, ax[u], bx[v]), sy = s1(hy, ay[u], by[v]), sz = s1(hz, az[u], bz[v]);
            prim[u * nb + v] = -0.5 * (d2(hx, ax[u], bx[v]) * sy * sz + sx * d2(hy, ay[u], by[v]) * sz +
                                       sx * sy * d2(hz, az[u], bz[v]));
          }
      } else {
        std::fill(prim.begin(), prim.end(), 0.0);
        const double px = (a * A.c[0] + b * B.c[0]) / p, py = (a * A.c[1] + b * B.c[1]) / p,
                     pz = (a * A.c[2] + b * B.c[2]) / p;
        for (int atom = 0; atom < natom; ++atom) {
          hermite_coulomb(la + lb, p, px - xyz[3 * atom], py - xyz[3 * atom + 1], pz - xyz[3 * atom + 2], rt);
          const double pref = -z[atom] * 2.0 * kPi / p;
          for (int u = 0; u < na; ++u)
            for (int v = 0; v < nb; ++v) {
              double s = 0.0;
              for (int t = 0; t <= ax[u] + bx[v]; ++t)
                for (int w = 0; w <= ay[u] + by[v]; ++w)
                  for (int q = 0; q <= az[u] + bz[v]; ++q)
                    s += hx.get(ax[u], bx[v], t) * hy.get(ay[u], by[v], w) * hz.get(az[u], bz[v], q) * rt.get(t, w, q);
              prim[u * nb + v] += pref * s;
            }
        }
      }
      for (int ca = 0; ca < A.ncontr; ++ca)
        for (int cb = 0; cb < B.ncontr; ++cb) {
          const double cc = A.coefs[(size_t)ca * npa + i] * B.coefs[(size_t)cb * npb + j];
          if (cc == 0.0) continue;
          for (int u = 0; u < na; ++u)
            for (int v = 0; v < nb; ++v) out[((size_t)ca * na + u) * (B.ncontr * nb) + cb * nb + v] += cc * prim[u * nb + v];
        }
    }
}

// Spherical transform of a contracted Cartesian pair block into the
// [nfunc_a][nfunc_b] block.
void to_sph_pair(const Shell& A, const Shell& B, const std::vector<double>& cart, std::vector<double>& sph) {
  const int la = A.l, lb = B.l, na = ncart(la), nb = ncart(lb), sa = 2 * la + 1, sb = 2 * lb + 1;
  const auto& ta = sph_table(la);
  const auto& tb = sph_table(lb);
  const int fa = A.nfunc(), fb = B.nfunc();
  sph.assign((size_t)fa * fb, 0.0);
  for (int ca = 0; ca < A.ncontr; ++ca)
    for (int cb = 0; cb < B.ncontr; ++cb)
      for (int ma = 0; ma < sa; ++ma)
        for (int mb = 0; mb < sb; ++mb) {
          double s = 0.0;
          for (int u = 0; u < na; ++u) {
            const double wa = ta[ma * na + u];
            if (wa == 0.0) continue;
            for (int v = 0; v < nb; ++v) {
              const double wb = tb[mb * nb + v];
              if (wb == 0.0) continue;
              s += wa * wb * cart[((size_t)ca * na + u) * (B.ncontr * nb) + cb * nb + v];
            }
          }
          sph[(size_t)(ca * sa + ma) * fb + cb * sb + mb] = s;
        }
}

// Norm factors 1/sqrt(<f|f>) of every function of a basis.
std::vector<double> norms(const std::vector<Shell>& shells, int nfunc) {
  std::vector<double> out(nfunc, 1.0);
  std::vector<double> cart, sph;
  for (const Shell& s : shells) {
    pair_1e(s, s, 0, 0, nullptr, nullptr, cart);
    to_sph_pair(s, s, cart, sph);
    const int f = s.nfunc();
    for (int k = 0; k < f; ++k) out[s.offset + k] = 1.0 / std::sqrt(sph[(size_t)k * f + k]);
  }
  return out;
}

// Coulomb Hermite contraction of (ab| c) over one primitive triple: the
// Cartesian block prim[(u * nb + v) * nc + w].
void coulomb3_prim(const Shell& A, const Shell& B, const Shell& C, double a, double b, double g,
                   const std::vector<int>& ax, const std::vector<int>& ay, const std::vector<int>& az,
                   const std::vector<int>& bx, const std::vector<int>& by, const std::vector<int>& bz,
                   const std::vector<int>& cx, const std::vector<int>& cy, const std::vector<int>& cz,
                   const Herm& hx, const Herm& hy, const Herm& hz, const Herm& gx, const Herm& gy, const Herm& gz,
                   std::vector<double>& prim) {
  const double p = a + b;
  const double px = (a * A.c[0] + b * B.c[0]) / p, py = (a * A.c[1] + b * B.c[1]) / p,
               pz = (a * A.c[2] + b * B.c[2]) / p;
  const double alpha = p * g / (p + g);
  RTab rt;
  const int L = A.l + B.l + C.l;
  hermite_coulomb(L, alpha, px - C.c[0], py - C.c[1], pz - C.c[2], rt);
  const double pref = 2.0 * std::pow(kPi, 2.5) / (p * g * std::sqrt(p + g));
  const int na = (int)ax.size(), nb = (int)bx.size(), nc = (int)cx.size();
  for (int u = 0; u < na; ++u)
    for (int v = 0; v < nb; ++v)
      for (int w = 0; w < nc; ++w) {
        double s = 0.0;
        for (int t1 = 0; t1 <= ax[u] + bx[v]; ++t1) {
          const double e1 = hx.get(ax[u], bx[v], t1);
          if (e1 == 0.0) continue;
          for (int u1 = 0; u1 <= ay[u] + by[v]; ++u1) {
            const double e2 = e1 * hy.get(ay[u], by[v], u1);
            if (e2 == 0.0) continue;
            for (int v1 = 0; v1 <= az[u] + bz[v]; ++v1) {
              const double e3 = e2 * hz.get(az[u], bz[v], v1);
              if (e3 == 0.0) continue;
              for (int t2 = 0; t2 <= cx[w]; ++t2) {
                const double f1 = gx.get(cx[w], 0, t2);
                if (f1 == 0.0) continue;
                for (int u2 = 0; u2 <= cy[w]; ++u2) {
                  const double f2 = f1 * gy.get(cy[w], 0, u2);
                  if (f2 == 0.0) continue;
                  for (int v2 = 0; v2 <= cz[w]; ++v2) {
                    const double f3 = f2 * gz.get(cz[w], 0, v2);
                    if (f3 == 0.0) continue;
                    const double sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                    s += e3 * f3 * sg * rt.get(t1 + t2, u1 + u2, v1 + v2);
                  }
                }
              }
            }
          }
        }
        prim[((size_t)u * nb + v) * nc + w] = pref * s;
      }
}

template <typename F>
void parallel_for(long long n, F&& fn) {
  unsigned nth = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
  if (n < 2 || nth == 1) {
    for (long long i = 0; i < n; ++i) fn(i);
    return;
  }
  std::vector<std::thread> pool;
  std::vector<std::string> errors(nth);
  for (unsigned w = 0; w < nth; ++w)
    pool.emplace_back([&, w] {
      try {
        for (long long i = w; i < n; i += nth) fn(i);
      } catch (const std::exception& e) {
        errors[w] = e.what();
      }
    });
  for (auto& t : pool) t.join();
  for (auto& e : errors)
    if (!e.empty()) throw std::runtime_error(e);
}

// (c | ab) for one aux shell and one orbital shell pair, spherical, normalised,
// written into out[(c_func) * n * n + a_func * n + b_func] (and the b, a mirror).
void triple(const Shell& A, const Shell& B, const Shell& C, const std::vector<double>& norm_o,
            const std::vector<double>& norm_x, int n, double* out) {
  std::vector<int> ax, ay, az, bx, by, bz, cx, cy, cz;
  cart_list(A.l, ax, ay, az);
  cart_list(B.l, bx, by, bz);
  cart_list(C.l, cx, cy, cz);
  const int na = ncart(A.l), nb = ncart(B.l), nc = ncart(C.l);
  const int npa = (int)A.exps.size(), npb = (int)B.exps.size(), npc = (int)C.exps.size();
  std::vector<double> cart((size_t)A.ncontr * na * B.ncontr * nb * C.ncontr * nc, 0.0);
  std::vector<double> prim((size_t)na * nb * nc);
  const double abx = A.c[0] - B.c[0], aby = A.c[1] - B.c[1], abz = A.c[2] - B.c[2];
  Herm hx, hy, hz, gx, gy, gz;
  for (int i = 0; i < npa; ++i)
    for (int j = 0; j < npb; ++j) {
      const double a = A.exps[i], b = B.exps[j];
      if (std::exp(-a * b / (a + b) * (abx * abx + aby * aby + abz * abz)) < kPairCut) continue;
      hermite(A.l, B.l, a, b, abx, hx);
      hermite(A.l, B.l, a, b, aby, hy);
      hermite(A.l, B.l, a, b, abz, hz);
      for (int k = 0; k < npc; ++k) {
        const double g = C.exps[k];
        hermite(C.l, 0, g, 0.0, 0.0, gx);
        hermite(C.l, 0, g, 0.0, 0.0, gy);
        hermite(C.l, 0, g, 0.0, 0.0, gz);
        coulomb3_prim(A, B, C, a, b, g, ax, ay, az, bx, by, bz, cx, cy, cz, hx, hy, hz, gx, gy, gz, prim);
        for (int ca = 0; ca < A.ncontr; ++ca)
          for (int cb = 0; cb < B.ncontr; ++cb)
            for (int cc = 0; cc < C.ncontr; ++cc) {
              const double w = A.coefs[(size_t)ca * npa + i] * B.coefs[(size_t)cb * npb + j] * C.coefs[(size_t)cc * npc + k];
              if (w == 0.0) continue;
              for (int u = 0; u < na; ++u)
                for (int v = 0; v < nb; ++v)
                  for (int q = 0; q < nc; ++q)
                    cart[((((size_t)ca * na + u) * B.ncontr + cb) * nb + v) * C.ncontr * nc + cc * nc + q] +=
                        w * prim[((size_t)u * nb + v) * nc + q];
            }
      }
    }
  // Spherical transform on all three indices.
  const auto& ta = sph_table(A.l);
  const auto& tb = sph_table(B.l);
  const auto& tc = sph_table(C.l);
  const int sa = 2 * A.l + 1, sb = 2 * B.l + 1, sc = 2 * C.l + 1;
  for (int ca = 0; ca < A.ncontr; ++ca)
    for (int cb = 0; cb < B.ncontr; ++cb)
      for (int cc = 0; cc < C.ncontr; ++cc)
        for (int ma = 0; ma < sa; ++ma)
          for (int mb = 0; mb < sb; ++mb)
            for (int mc = 0; mc < sc; ++mc) {
              double s = 0.0;
              for (int u = 0; u < na; ++u) {
                const double wa = ta[ma * na + u];
                if (wa == 0.0) continue;
                for (int v = 0; v < nb; ++v) {
                  const double wb = tb[mb * nb + v];
                  if (wb == 0.0) continue;
                  for (int q = 0; q < nc; ++q) {
                    const double wc = tc[mc * nc + q];
                    if (wc == 0.0) continue;
                    s += wa * wb * wc * cart[((((size_t)ca * na + u) * B.ncontr + cb) * nb + v) * C.ncontr * nc + cc * nc + q];
                  }
                }
              }
              const int fa = A.offset + ca * sa + ma, fb = B.offset + cb * sb + mb, fc = C.offset + cc * sc + mc;
              const double val = s * norm_o[fa] * norm_o[fb] * norm_x[fc];
              out[(size_t)fc * n * n + (size_t)fa * n + fb] = val;
              out[(size_t)fc * n * n + (size_t)fb * n + fa] = val;
            }
}

// Spherical transform of one axis of a 4-index block: dims[axis] goes from
// S.ncontr * ncart(S.l) Cartesian to S.ncontr * (2 l + 1) spherical entries.
void sph_axis(const std::vector<double>& in, int dims[4], int axis, const Shell& S, std::vector<double>& out) {
  const int nc = ncart(S.l), ns = 2 * S.l + 1;
  const auto& t = sph_table(S.l);
  long long inner = 1, outer = 1;
  for (int k = axis + 1; k < 4; ++k) inner *= dims[k];
  for (int k = 0; k < axis; ++k) outer *= dims[k];
  const int nout = S.ncontr * ns;
  out.assign((size_t)(outer * nout * inner), 0.0);
  for (long long o = 0; o < outer; ++o)
    for (int c = 0; c < S.ncontr; ++c)
      for (int m = 0; m < ns; ++m) {
        double* dst = &out[(size_t)((o * nout + c * ns + m) * inner)];
        for (int u = 0; u < nc; ++u) {
          const double w = t[m * nc + u];
          if (w == 0.0) continue;
          const double* src = &in[(size_t)((o * dims[axis] + c * nc + u) * inner)];
          for (long long i = 0; i < inner; ++i) dst[i] += w * src[i];
        }
      }
  dims[axis] = nout;
}

// (ab|cd) for one shell quartet, spherical and normalised, written into
// out[((a * n + b) * n + c) * n + d] for all eight index permutations.
void quartet(const Shell& A, const Shell& B, const Shell& C, const Shell& D, const std::vector<double>& norm, int n,
             double* out) {
  const Shell* S[4] = {&A, &B, &C, &D};
  std::vector<int> cx[4], cy[4], cz[4];
  int nc[4], np[4];
  for (int k = 0; k < 4; ++k) {
    cart_list(S[k]->l, cx[k], cy[k], cz[k]);
    nc[k] = ncart(S[k]->l);
    np[k] = (int)S[k]->exps.size();
  }
  int dims[4];
  for (int k = 0; k < 4; ++k) dims[k] = S[k]->ncontr * nc[k];
  std::vector<double> cart((size_t)dims[0] * dims[1] * dims[2] * dims[3], 0.0);
  std::vector<double> prim((size_t)nc[0] * nc[1] * nc[2] * nc[3]);
  double ab[3], cd[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = A.c[k] - B.c[k];
    cd[k] = C.c[k] - D.c[k];
  }
  const double ab2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
  const double cd2 = cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2];
  const int L = A.l + B.l + C.l + D.l;
  Herm h[3], g[3];
  RTab rt;
  for (int i = 0; i < np[0]; ++i)
    for (int j = 0; j < np[1]; ++j) {
      const double a = A.exps[i], b = B.exps[j], p = a + b;
      if (std::exp(-a * b / p * ab2) < kPairCut) continue;
      for (int k = 0; k < 3; ++k) hermite(A.l, B.l, a, b, ab[k], h[k]);
      double P[3];
      for (int k = 0; k < 3; ++k) P[k] = (a * A.c[k] + b * B.c[k]) / p;
      for (int k2 = 0; k2 < np[2]; ++k2)
        for (int l2 = 0; l2 < np[3]; ++l2) {
          const double c = C.exps[k2], d = D.exps[l2], q = c + d;
          if (std::exp(-c * d / q * cd2) < kPairCut) continue;
          for (int k = 0; k < 3; ++k) hermite(C.l, D.l, c, d, cd[k], g[k]);
          double Q[3];
          for (int k = 0; k < 3; ++k) Q[k] = (c * C.c[k] + d * D.c[k]) / q;
          hermite_coulomb(L, p * q / (p + q), P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], rt);
          const double pref = 2.0 * std::pow(kPi, 2.5) / (p * q * std::sqrt(p + q));
          for (int u = 0; u < nc[0]; ++u)
            for (int v = 0; v < nc[1]; ++v)
              for (int w = 0; w < nc[2]; ++w)
                for (int x = 0; x < nc[3]; ++x) {
                  double sum = 0.0;
                  for (int t1 = 0; t1 <= cx[0][u] + cx[1][v]; ++t1) {
                    const double e1 = h[0].get(cx[0][u], cx[1][v], t1);
                    if (e1 == 0.0) continue;
                    for (int u1 = 0; u1 <= cy[0][u] + cy[1][v]; ++u1) {
                      const double e2 = e1 * h[1].get(cy[0][u], cy[1][v], u1);
                      if (e2 == 0.0) continue;
                      for (int v1 = 0; v1 <= cz[0][u] + cz[1][v]; ++v1) {
                        const double e3 = e2 * h[2].get(cz[0][u], cz[1][v], v1);
                        if (e3 == 0.0) continue;
                        for (int t2 = 0; t2 <= cx[2][w] + cx[3][x]; ++t2) {
                          const double f1 = g[0].get(cx[2][w], cx[3][x], t2);
                          if (f1 == 0.0) continue;
                          for (int u2 = 0; u2 <= cy[2][w] + cy[3][x]; ++u2) {
                            const double f2 = f1 * g[1].get(cy[2][w], cy[3][x], u2);
                            if (f2 == 0.0) continue;
                            for (int v2 = 0; v2 <= cz[2][w] + cz[3][x]; ++v2) {
                              const double f3 = f2 * g[2].get(cz[2][w], cz[3][x], v2);
                              if (f3 == 0.0) continue;
                              const double sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                              sum += e3 * f3 * sg * rt.get(t1 + t2, u1 + u2, v1 + v2);
                            }
                          }
                        }
                      }
                    }
                  }
                  prim[(((size_t)u * nc[1] + v) * nc[2] + w) * nc[3] + x] = pref * sum;
                }
          for (int ca = 0; ca < A.ncontr; ++ca)
            for (int cb = 0; cb < B.ncontr; ++cb)
              for (int cc = 0; cc < C.ncontr; ++cc)
                for (int cdd = 0; cdd < D.ncontr; ++cdd) {
                  const double wgt = A.coefs[(size_t)ca * np[0] + i] * B.coefs[(size_t)cb * np[1] + j] *
                                     C.coefs[(size_t)cc * np[2] + k2] * D.coefs[(size_t)cdd * np[3] + l2];
                  if (wgt == 0.0) continue;
                  for (int u = 0; u < nc[0]; ++u)
                    for (int v = 0; v < nc[1]; ++v)
                      for (int w = 0; w < nc[2]; ++w)
                        for (int x = 0; x < nc[3]; ++x)
                          cart[(((size_t)(ca * nc[0] + u) * dims[1] + cb * nc[1] + v) * dims[2] + cc * nc[2] + w) *
                                   dims[3] +
                               cdd * nc[3] + x] += wgt * prim[(((size_t)u * nc[1] + v) * nc[2] + w) * nc[3] + x];
                }
        }
    }
  std::vector<double> tmp;
  for (int k = 0; k < 4; ++k) {
    sph_axis(cart, dims, k, *S[k], tmp);
    cart.swap(tmp);
  }
  for (int fa = 0; fa < dims[0]; ++fa)
    for (int fb = 0; fb < dims[1]; ++fb)
      for (int fc = 0; fc < dims[2]; ++fc)
        for (int fd = 0; fd < dims[3]; ++fd) {
          const int a = A.offset + fa, b = B.offset + fb, c = C.offset + fc, d = D.offset + fd;
          const double val = cart[(((size_t)fa * dims[1] + fb) * dims[2] + fc) * dims[3] + fd] * norm[a] * norm[b] *
                             norm[c] * norm[d];
          const size_t N = (size_t)n;
          const int idx[8][4] = {{a, b, c, d}, {b, a, c, d}, {a, b, d, c}, {b, a, d, c},
                                 {c, d, a, b}, {d, c, a, b}, {c, d, b, a}, {d, c, b, a}};
          for (const auto& q : idx) out[((q[0] * N + q[1]) * N + q[2]) * N + q[3]] = val;
        }
}

template <typename F>
int iguard(F&& body) {
  try {
    for (int l = 0; l <= kMaxL; ++l) sph_table(l);  // build the shared tables before any thread reads them
    body();
    return HFJK_OK;
  } catch (const std::invalid_argument& e) {
    hfjk_internal_set_error(e.what());
    return HFJK_ERR_INPUT;
  } catch (const std::exception& e) {
    hfjk_internal_set_error(e.what());
    return HFJK_ERR_HIP;
  }
}

}  // namespace

extern "C" {

int hfjk_basis_nfunc(const hfjk_basis* basis, int* nfunc) {
  return iguard([&] {
    if (nfunc == nullptr) throw std::invalid_argument("null output pointer");
    make_shells(basis, *nfunc);
  });
}

int hfjk_int1e(const hfjk_basis* basis, int natom, const double* charges, const double* coords, double* overlap,
               double* kinetic, double* nuclear) {
  return iguard([&] {
    int n = 0;
    const std::vector<Shell> sh = make_shells(basis, n);
    if (natom < 0 || (natom > 0 && (charges == nullptr || coords == nullptr)))
      throw std::invalid_argument("nuclear attraction needs charges and coordinates");
    const std::vector<double> nm = norms(sh, n);
    const long long ns = (long long)sh.size();
    const long long npair = ns * (ns + 1) / 2;
    parallel_for(npair, [&](long long idx) {
      long long s1 = (long long)((std::sqrt(8.0 * idx + 1.0) - 1.0) / 2.0);
      while (s1 * (s1 + 1) / 2 > idx) --s1;
      while ((s1 + 1) * (s1 + 2) / 2 <= idx) ++s1;
      const long long s2 = idx - s1 * (s1 + 1) / 2;
      const Shell& A = sh[s1];
      const Shell& B = sh[s2];
      std::vector<double> cart, sph;
      double* outs[3] = {overlap, kinetic, nuclear};
      for (int kind = 0; kind < 3; ++kind) {
        if (outs[kind] == nullptr) continue;
        pair_1e(A, B, kind, natom, charges, coords, cart);
        to_sph_pair(A, B, cart, sph);
        const int fa = A.nfunc(), fb = B.nfunc();
        for (int u = 0; u < fa; ++u)
          for (int v = 0; v < fb; ++v) {
            const int i = A.offset + u, j = B.offset + v;
            const double val = sph[(size_t)u * fb + v] * nm[i] * nm[j];
            outs[kind][(size_t)i * n + j] = val;
            outs[kind][(size_t)j * n + i] = val;
          }
      }
    });
  });
}

int hfjk_int2c2e(const hfjk_basis* aux, double* metric) {
  return iguard([&] {
    int n = 0;
    const std::vector<Shell> sh = make_shells(aux, n);
    if (metric == nullptr) throw std::invalid_argument("null output pointer");
    const std::vector<double> nm = norms(sh, n);
    // (P|Q) = (P s0 | Q) with a unit s function of exponent 0 on P's centre.
    const long long ns = (long long)sh.size();
    const long long npair = ns * (ns + 1) / 2;
    Shell unit;
    unit.l = 0;
    unit.exps = {0.0};
    unit.coefs = {1.0};
    unit.ncontr = 1;
    unit.offset = 0;
    parallel_for(npair, [&](long long idx) {
      long long s1 = (long long)((std::sqrt(8.0 * idx + 1.0) - 1.0) / 2.0);
      while (s1 * (s1 + 1) / 2 > idx) --s1;
      while ((s1 + 1) * (s1 + 2) / 2 <= idx) ++s1;
      const long long s2 = idx - s1 * (s1 + 1) / 2;
      Shell u = unit;
      std::memcpy(u.c, sh[s1].c, sizeof(u.c));
      // Evaluate (P 1 | Q) as a three-centre block with a scratch layout of
      // one function for the unit shell.
      const Shell& P = sh[s1];
      const Shell& Q = sh[s2];
      const int fp = P.nfunc(), fq = Q.nfunc();
      const int m = std::max(fp, 1);
      std::vector<double> block((size_t)fq * m * m, 0.0);
      Shell Pl = P;
      Pl.offset = 0;
      Shell ul = u;
      ul.offset = 0;
      Shell Ql = Q;
      Ql.offset = 0;
      std::vector<double> ones_p(fp, 1.0), ones_q(fq, 1.0);
      // triple() writes out[fc][fa][fb] and its mirror; with B = unit (one
      // function, index 0) the values sit at [fc][fa][0].
      std::vector<double> tmp((size_t)fq * fp * fp, 0.0);
      triple(Pl, ul, Ql, ones_p, ones_q, fp, tmp.data());
      for (int a = 0; a < fp; ++a)
        for (int q = 0; q < fq; ++q) {
          const int i = P.offset + a, j = Q.offset + q;
          const double val = tmp[(size_t)q * fp * fp + (size_t)a * fp + 0] * nm[i] * nm[j];
          metric[(size_t)i * n + j] = val;
          metric[(size_t)j * n + i] = val;
        }
    });
  });
}

int hfjk_int2e(const hfjk_basis* basis, double* out) {
  return iguard([&] {
    int n = 0;
    const std::vector<Shell> sh = make_shells(basis, n);
    if (out == nullptr) throw std::invalid_argument("null output pointer");
    const std::vector<double> nm = norms(sh, n);
    // Unique quartets: shell pairs (s1 >= s2) and pair index p12 >= p34.
    const long long ns = (long long)sh.size();
    const long long npair = ns * (ns + 1) / 2;
    std::vector<std::pair<int, int>> pairs;
    for (int s1 = 0; s1 < ns; ++s1)
      for (int s2 = 0; s2 <= s1; ++s2) pairs.emplace_back(s1, s2);
    const long long total = npair * (npair + 1) / 2;
    parallel_for(total, [&](long long idx) {
      long long p12 = (long long)((std::sqrt(8.0 * idx + 1.0) - 1.0) / 2.0);
      while (p12 * (p12 + 1) / 2 > idx) --p12;
      while ((p12 + 1) * (p12 + 2) / 2 <= idx) ++p12;
      const long long p34 = idx - p12 * (p12 + 1) / 2;
      quartet(sh[pairs[p12].first], sh[pairs[p12].second], sh[pairs[p34].first], sh[pairs[p34].second], nm, n, out);
    });
  });
}

int hfjk_int3c2e(const hfjk_basis* aux, const hfjk_basis* orbital, double* out) {
  return iguard([&] {
    int nx = 0, n = 0;
    const std::vector<Shell> sx = make_shells(aux, nx);
    const std::vector<Shell> so = make_shells(orbital, n);
    if (out == nullptr) throw std::invalid_argument("null output pointer");
    const std::vector<double> nmx = norms(sx, nx);
    const std::vector<double> nmo = norms(so, n);
    const long long ns = (long long)so.size();
    const long long npair = ns * (ns + 1) / 2;
    const long long total = npair * (long long)sx.size();
    parallel_for(total, [&](long long idx) {
      const long long pair = idx % npair;
      const long long c = idx / npair;
      long long s1 = (long long)((std::sqrt(8.0 * pair + 1.0) - 1.0) / 2.0);
      while (s1 * (s1 + 1) / 2 > pair) --s1;
      while ((s1 + 1) * (s1 + 2) / 2 <= pair) ++s1;
      const long long s2 = pair - s1 * (s1 + 1) / 2;
      triple(so[s1], so[s2], sx[c], nmo, nmx, n, out);
    });
  });
}

}  // extern "C"

// ------------------------------------------------------------------ on the GPU
//
// (P|uv) and (P|Q) on the device: one 64-lane workgroup per shell triple
// (aux shell C, orbital shell pair A >= B; for the metric A is an aux shell
// and B a unit s function of exponent 0 on A's centre). Per primitive triple
// the wave forms, in LDS, the Hermite expansion coefficients of the A B pair
// and of C along x, y, z and the Hermite Coulomb integrals R_tuv (Boys
// function, then the downward recursion one order per step, the lanes
// sharing the entries of an order); each lane then contracts the Cartesian
// components it owns and accumulates them, weighted by the contraction
// coefficients, into the shell triple's Cartesian block in LDS. Last, the
// lanes apply the spherical transforms and the normalisation and write the
// block (and its (u, v) mirror). Shell pairs whose every primitive pair is
// below the product cut are skipped on the host, as the host engine skips
// them primitive by primitive; their entries stay zero.
namespace {

constexpr int kGL = 5;                  // largest l of a GPU shell (h)
constexpr int kGMaxL = 12;              // largest total angular momentum of a triple
constexpr int kGRtab = (kGMaxL + 1) * (kGMaxL + 2) * (kGMaxL + 3) / 6;
constexpr int kGCart = 1024;            // Cartesian block entries per triple
constexpr int kGEA = 128;               // Hermite coefficients of a pair per dimension
constexpr int kGEC = (kGL + 1) * (kGL + 1);  // ... of a single shell (its partner index fixed at 0)

struct DShell {
  int l, nprim, ncontr, eoff, coff, foff;
  double c[3];
};

struct DItem {
  int a, b, c;  // b < 0: the unit s function on a's centre
};

__device__ __forceinline__ int ncart_d(int l) { return (l + 1) * (l + 2) / 2; }

// Index of (t, u, v), t + u + v <= L, in a table ordered by the sum s, then u + v, then v.
__device__ __forceinline__ int rindex(int t, int u, int v) {
  const int s = t + u + v, m = u + v;
  return s * (s + 1) * (s + 2) / 6 + m * (m + 1) / 2 + v;
}

// Cartesian component k of l (x power descending, then y): (px, py, pz).
__device__ __forceinline__ void cart_of(int l, int k, int& px, int& py, int& pz) {
  int i = l, kk = k;
  while (kk > l - i) {
    kk -= l - i + 1;
    --i;
  }
  px = i;
  py = l - i - kk;
  pz = kk;
}

// The Boys function on the GPU: below kBoysTMax from a table of F_n on a
// grid of step kBoysStep (n <= kGMaxL + kBoysTaylor), F_nmax by a Taylor
// expansion about the nearest grid point (dF_n/dT = -F_{n+1}; the remainder
// is below 1e-15 relative for |dT| <= kBoysStep / 2), then the stable
// downward recursion; above it the erf closed form and upward recursion.
constexpr double kBoysStep = 0.05;
constexpr double kBoysTMax = 45.0;
constexpr int kBoysTaylor = 7;
constexpr int kBoysOrders = kGMaxL + kBoysTaylor + 1;
constexpr int kBoysPoints = (int)(kBoysTMax / kBoysStep) + 2;

__device__ void boys_d(int nmax, double t, const double* __restrict__ table, double* f) {
  const double et = exp(-t);
  if (t < kBoysTMax) {
    const int g = (int)(t / kBoysStep + 0.5);
    const double dt = t - g * kBoysStep;
    const double* row = table + (long long)g * kBoysOrders + nmax;
    double term = 1.0, sum = row[0];
    for (int k = 1; k < kBoysTaylor; ++k) {
      term *= -dt / k;
      sum += term * row[k];
    }
    f[nmax] = sum;
    for (int n = nmax - 1; n >= 0; --n) f[n] = (2.0 * t * f[n + 1] + et) / (2 * n + 1);
  } else {
    f[0] = 0.5 * sqrt(kPi / t) * erf(sqrt(t));
    for (int n = 0; n < nmax; ++n) f[n + 1] = ((2 * n + 1) * f[n] - et) / (2.0 * t);
  }
}

// The table of boys_d: F_n(g kBoysStep) for n < kBoysOrders, from the host
// evaluation (series and downward recursion).
std::vector<double> boys_table() {
  std::vector<double> out((size_t)kBoysPoints * kBoysOrders);
  for (int g = 0; g < kBoysPoints; ++g) boys(kBoysOrders - 1, g * kBoysStep, &out[(size_t)g * kBoysOrders]);
  return out;
}

// One dimension of the Hermite expansion of a primitive pair: e[(i (lb + 1) + j) nt + t].
__device__ void hermite_d(int la, int lb, double a, double b, double xab, double* e) {
  const int nt = la + lb + 1;
  for (int k = 0; k < (la + 1) * (lb + 1) * nt; ++k) e[k] = 0.0;
  const double p = a + b, mu = p > 0.0 ? a * b / p : 0.0;
  const double xpa = p > 0.0 ? -b / p * xab : 0.0, xpb = p > 0.0 ? a / p * xab : 0.0, i2p = 0.5 / p;
  auto get = [&](int i, int j, int t) { return (t < 0 || t > i + j) ? 0.0 : e[(i * (lb + 1) + j) * nt + t]; };
  e[0] = exp(-mu * xab * xab);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      for (int t = 0; t <= i + j; ++t) {
        double v;
        if (i > 0)
          v = i2p * get(i - 1, j, t - 1) + xpa * get(i - 1, j, t) + (t + 1) * get(i - 1, j, t + 1);
        else
          v = i2p * get(i, j - 1, t - 1) + xpb * get(i, j - 1, t) + (t + 1) * get(i, j - 1, t + 1);
        e[(i * (lb + 1) + j) * nt + t] = v;
      }
    }
}

struct TripleSmem {
  double ea[3][kGEA];       // A B pair, per dimension
  double ec[3][kGEC];       // C alone (its B index fixed at 0)
  double r[2][kGRtab];      // Hermite Coulomb integrals, two orders
  double f[kGMaxL + 1];     // Boys values
  double cart[kGCart];      // contracted Cartesian block
};

__global__ __launch_bounds__(64) void triple_kernel(const DItem* __restrict__ items, const DShell* __restrict__ orb,
                                                    const DShell* __restrict__ aux, const double* __restrict__ exps,
                                                    const double* __restrict__ coefs,
                                                    const double* __restrict__ norm_o, const double* __restrict__ norm_x,
                                                    const double* __restrict__ sph, const int* __restrict__ sph_off,
                                                    const double* __restrict__ boys_tab, int n, int metric_mode,
                                                    double* __restrict__ out) {
  __shared__ TripleSmem sm;
  const int lane = threadIdx.x;
  const DItem it = items[blockIdx.x];
  // Shells: A and B from the orbital list (three-centre) or A from the aux list and B the unit
  // function (metric); C from the aux list.
  const DShell A = metric_mode ? aux[it.a] : orb[it.a];
  DShell B;
  if (it.b >= 0) {
    B = orb[it.b];
  } else {
    B.l = 0;
    B.nprim = 1;
    B.ncontr = 1;
    B.eoff = -1;
    B.coff = -1;
    B.foff = 0;
    B.c[0] = A.c[0];
    B.c[1] = A.c[1];
    B.c[2] = A.c[2];
  }
  const DShell C = aux[it.c];
  const int na = ncart_d(A.l), nb = ncart_d(B.l), nc = ncart_d(C.l);
  const int da = A.ncontr * na, db = B.ncontr * nb, dc = C.ncontr * nc;
  const int ncomp = na * nb * nc;
  const int L = A.l + B.l + C.l;
  const int nta = A.l + B.l + 1, ntc = C.l + 1;
  for (int k = lane; k < da * db * dc; k += 64) sm.cart[k] = 0.0;
  const double abx = A.c[0] - B.c[0], aby = A.c[1] - B.c[1], abz = A.c[2] - B.c[2];
  const double ab2 = abx * abx + aby * aby + abz * abz;
  for (int i = 0; i < A.nprim; ++i)
    for (int j = 0; j < B.nprim; ++j) {
      const double a = exps[A.eoff + i];
      const double b = B.eoff >= 0 ? exps[B.eoff + j] : 0.0;
      const double p = a + b;
      if (exp(-a * b / p * ab2) < kPairCut) continue;
      const double px = (a * A.c[0] + b * B.c[0]) / p, py = (a * A.c[1] + b * B.c[1]) / p,
                   pz = (a * A.c[2] + b * B.c[2]) / p;
      __syncthreads();  // the previous pair's tables are no longer read
      if (lane < 3) hermite_d(A.l, B.l, a, b, lane == 0 ? abx : lane == 1 ? aby : abz, sm.ea[lane]);
      for (int k = 0; k < C.nprim; ++k) {
        const double g = exps[C.eoff + k];
        const double alpha = p * g / (p + g);
        const double xq = px - C.c[0], yq = py - C.c[1], zq = pz - C.c[2];
        __syncthreads();  // the previous triple's R and E_C are no longer read
        if (lane >= 3 && lane < 6) hermite_d(C.l, 0, g, 0.0, 0.0, sm.ec[lane - 3]);
        if (lane == 6) boys_d(L, alpha * (xq * xq + yq * yq + zq * zq), boys_tab, sm.f);
        __syncthreads();
        // R^(L)_000 .. R^(0)_tuv: order m holds the entries with t + u + v <= L - m.
        int cur = 0;
        if (lane == 0) sm.r[0][0] = pow(-2.0 * alpha, L) * sm.f[L];
        for (int m = L - 1; m >= 0; --m) {
          __syncthreads();
          const int top = L - m;
          const int count = (top + 1) * (top + 2) * (top + 3) / 6;
          const double fac = pow(-2.0 * alpha, m);
          double* dst = sm.r[cur ^ 1];
          const double* src = sm.r[cur];
          for (int idx = lane; idx < count; idx += 64) {
            // Decode idx -> (t, u, v) in the rindex order.
            int s = 0;
            while ((s + 1) * (s + 2) * (s + 3) / 6 <= idx) ++s;
            int rem = idx - s * (s + 1) * (s + 2) / 6;
            int mm = 0;
            while ((mm + 1) * (mm + 2) / 2 <= rem) ++mm;
            const int v = rem - mm * (mm + 1) / 2, u = mm - v, t = s - mm;
            double val;
            if (s == 0)
              val = fac * sm.f[m];
            else if (t > 0)
              val = (t > 1 ? (t - 1) * src[rindex(t - 2, u, v)] : 0.0) + xq * src[rindex(t - 1, u, v)];
            else if (u > 0)
              val = (u > 1 ? (u - 1) * src[rindex(t, u - 2, v)] : 0.0) + yq * src[rindex(t, u - 1, v)];
            else
              val = (v > 1 ? (v - 1) * src[rindex(t, u, v - 2)] : 0.0) + zq * src[rindex(t, u, v - 1)];
            dst[idx] = val;
          }
          cur ^= 1;
        }
        __syncthreads();
        const double* R = sm.r[cur];
        const double pref = 2.0 * pow(kPi, 2.5) / (p * g * sqrt(p + g));
        for (int q = lane; q < ncomp; q += 64) {
          const int ia = q / (nb * nc), ib = (q / nc) % nb, ic = q % nc;
          int ax, ay, az, bx, by, bz, cx, cy, cz;
          cart_of(A.l, ia, ax, ay, az);
          cart_of(B.l, ib, bx, by, bz);
          cart_of(C.l, ic, cx, cy, cz);
          const double* ex = &sm.ea[0][(ax * (B.l + 1) + bx) * nta];
          const double* ey = &sm.ea[1][(ay * (B.l + 1) + by) * nta];
          const double* ez = &sm.ea[2][(az * (B.l + 1) + bz) * nta];
          const double* gx = &sm.ec[0][cx * ntc];
          const double* gy = &sm.ec[1][cy * ntc];
          const double* gz = &sm.ec[2][cz * ntc];
          double sum = 0.0;
          for (int t1 = 0; t1 <= ax + bx; ++t1) {
            const double e1 = ex[t1];
            if (e1 == 0.0) continue;
            for (int u1 = 0; u1 <= ay + by; ++u1) {
              const double e2 = e1 * ey[u1];
              if (e2 == 0.0) continue;
              for (int v1 = 0; v1 <= az + bz; ++v1) {
                const double e3 = e2 * ez[v1];
                if (e3 == 0.0) continue;
                for (int t2 = 0; t2 <= cx; ++t2) {
                  const double f1 = gx[t2];
                  if (f1 == 0.0) continue;
                  for (int u2 = 0; u2 <= cy; ++u2) {
                    const double f2 = f1 * gy[u2];
                    if (f2 == 0.0) continue;
                    for (int v2 = 0; v2 <= cz; ++v2) {
                      const double f3 = f2 * gz[v2];
                      if (f3 == 0.0) continue;
                      const double sgn = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                      sum += e3 * f3 * sgn * R[rindex(t1 + t2, u1 + u2, v1 + v2)];
                    }
                  }
                }
              }
            }
          }
          const double prim = pref * sum;
          for (int ca = 0; ca < A.ncontr; ++ca) {
            const double wa = coefs[A.coff + ca * A.nprim + i];
            if (wa == 0.0) continue;
            for (int cb = 0; cb < B.ncontr; ++cb) {
              const double wb = B.coff >= 0 ? coefs[B.coff + cb * B.nprim + j] : 1.0;
              if (wb == 0.0) continue;
              for (int cc = 0; cc < C.ncontr; ++cc) {
                const double wc = coefs[C.coff + cc * C.nprim + k];
                if (wc == 0.0) continue;
                sm.cart[((ca * na + ia) * db + cb * nb + ib) * dc + cc * nc + ic] += wa * wb * wc * prim;
              }
            }
          }
        }
      }
    }
  __syncthreads();
  // Spherical transform of the three indices, normalisation, and the stores.
  const int sa = 2 * A.l + 1, sbs = 2 * B.l + 1, sc = 2 * C.l + 1;
  const int fa_n = A.ncontr * sa, fb_n = B.ncontr * sbs, fc_n = C.ncontr * sc;
  const double* ta = sph + sph_off[A.l];
  const double* tb = sph + sph_off[B.l];
  const double* tc = sph + sph_off[C.l];
  for (int o = lane; o < fa_n * fb_n * fc_n; o += 64) {
    const int fa = o / (fb_n * fc_n), fb = (o / fc_n) % fb_n, fc = o % fc_n;
    const int ca = fa / sa, ma = fa % sa, cb = fb / sbs, mb = fb % sbs, cc = fc / sc, mc = fc % sc;
    double s = 0.0;
    for (int u = 0; u < na; ++u) {
      const double wa = ta[ma * na + u];
      if (wa == 0.0) continue;
      for (int v = 0; v < nb; ++v) {
        const double wb = tb[mb * nb + v];
        if (wb == 0.0) continue;
        const double* row = &sm.cart[((ca * na + u) * db + cb * nb + v) * dc + cc * nc];
        for (int w = 0; w < nc; ++w) {
          const double wc = tc[mc * nc + w];
          if (wc != 0.0) s += wa * wb * wc * row[w];
        }
      }
    }
    if (metric_mode) {
      const long long i = A.foff + fa, jj = C.foff + fc;
      const double val = s * norm_x[i] * norm_x[jj];
      out[i * n + jj] = val;
      out[jj * n + i] = val;
    } else {
      const long long ia = A.foff + fa, ib = B.foff + fb, ic = C.foff + fc;
      const double val = s * norm_o[ia] * norm_o[ib] * norm_x[ic];
      out[(ic * n + ia) * n + ib] = val;
      out[(ic * n + ib) * n + ia] = val;
    }
  }
}

// Device copies of the shell tables of one or two bases, the spherical
// tables, the norms and the item list; freed with the object.
struct DeviceTables {
  std::vector<void*> buffers;
  ~DeviceTables() {
    for (void* p : buffers) (void)hipFree(p);
  }
  template <typename T>
  T* upload(const std::vector<T>& host, hipStream_t s) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(host.size(), 1) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc of the integral tables failed");
    buffers.push_back(p);
    if (!host.empty() && hipMemcpyAsync(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
      throw std::runtime_error("hipMemcpyAsync of the integral tables failed");
    return (T*)p;
  }
};

void device_shells(const std::vector<Shell>& shells, std::vector<DShell>& out, std::vector<double>& exps,
                   std::vector<double>& coefs) {
  for (const Shell& s : shells) {
    if (s.l > kGL) throw std::invalid_argument("the GPU integrals take l <= 5");
    DShell d;
    d.l = s.l;
    d.nprim = (int)s.exps.size();
    d.ncontr = s.ncontr;
    d.eoff = (int)exps.size();
    d.coff = (int)coefs.size();
    d.foff = s.offset;
    for (int k = 0; k < 3; ++k) d.c[k] = s.c[k];
    exps.insert(exps.end(), s.exps.begin(), s.exps.end());
    coefs.insert(coefs.end(), s.coefs.begin(), s.coefs.end());
    out.push_back(d);
  }
}

// Whether any primitive pair of two shells passes the product cut (the host engine's per-pair test).
bool pair_survives(const Shell& A, const Shell& B) {
  const double d2 = (A.c[0] - B.c[0]) * (A.c[0] - B.c[0]) + (A.c[1] - B.c[1]) * (A.c[1] - B.c[1]) +
                    (A.c[2] - B.c[2]) * (A.c[2] - B.c[2]);
  for (double a : A.exps)
    for (double b : B.exps)
      if (std::exp(-a * b / (a + b) * d2) >= kPairCut) return true;
  return false;
}

size_t cart_entries(const Shell& A, int bl, int bcontr, const Shell& C) {
  return (size_t)A.ncontr * ncart(A.l) * bcontr * ncart(bl) * C.ncontr * ncart(C.l);
}

void launch_triples(const std::vector<Shell>& orb, const std::vector<Shell>& aux, const std::vector<DItem>& items,
                    const std::vector<double>& norm_o, const std::vector<double>& norm_x, int n, bool metric_mode,
                    double* out, hipStream_t s) {
  std::vector<DShell> dorb, daux;
  std::vector<double> exps, coefs;
  device_shells(orb, dorb, exps, coefs);
  device_shells(aux, daux, exps, coefs);
  std::vector<double> sph;
  std::vector<int> sph_off;
  for (int l = 0; l <= kGL; ++l) {
    sph_off.push_back((int)sph.size());
    const auto& t = sph_table(l);
    sph.insert(sph.end(), t.begin(), t.end());
  }
  DeviceTables dev;
  const DItem* d_items = dev.upload(items, s);
  const DShell* d_orb = dev.upload(dorb, s);
  const DShell* d_aux = dev.upload(daux, s);
  const double* d_exps = dev.upload(exps, s);
  const double* d_coefs = dev.upload(coefs, s);
  const double* d_no = dev.upload(norm_o, s);
  const double* d_nx = dev.upload(norm_x, s);
  const double* d_sph = dev.upload(sph, s);
  const int* d_off = dev.upload(sph_off, s);
  const std::vector<double> boys_host = boys_table();  // alive until the synchronisation below
  const double* d_boys = dev.upload(boys_host, s);
  const size_t total = metric_mode ? (size_t)n * n : (size_t)norm_x.size() * n * n;
  if (hipMemsetAsync(out, 0, total * sizeof(double), s) != hipSuccess) throw std::runtime_error("hipMemsetAsync failed");
  // At most 2^31 - 1 workgroups per launch.
  const size_t per = (size_t)1 << 30;
  for (size_t base = 0; base < items.size(); base += per) {
    const unsigned count = (unsigned)std::min(per, items.size() - base);
    hipLaunchKernelGGL(triple_kernel, dim3(count), dim3(64), 0, s, d_items + base, d_orb, d_aux, d_exps, d_coefs,
                       d_no, d_nx, d_sph, d_off, d_boys, n, metric_mode ? 1 : 0, out);
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("launch of the integral kernel failed");
  }
  // The tables are freed on return: wait for the kernels that read them.
  if (hipStreamSynchronize(s) != hipSuccess) throw std::runtime_error("the integral kernel failed");
}

void check_item_sizes(const std::vector<Shell>& a_list, const std::vector<Shell>& c_list, bool unit_b,
                      const std::vector<Shell>* b_list) {
  int la = 0, lb = 0, lc = 0;
  size_t worst = 0;
  for (const Shell& s : a_list) la = std::max(la, s.l);
  for (const Shell& s : c_list) lc = std::max(lc, s.l);
  if (b_list)
    for (const Shell& s : *b_list) lb = std::max(lb, s.l);
  if (la + lb + lc > kGMaxL) throw std::invalid_argument("total angular momentum beyond 12 on the GPU");
  if ((la + 1) * (lb + 1) * (la + lb + 1) > kGEA)
    throw std::invalid_argument("a shell pair's Hermite table exceeds the GPU kernel's LDS table");
  for (const Shell& A : a_list)
    for (const Shell& C : c_list) {
      if (unit_b) worst = std::max(worst, cart_entries(A, 0, 1, C));
      else
        for (const Shell& B : *b_list) worst = std::max(worst, cart_entries(A, B.l, B.ncontr, C));
    }
  if (worst > (size_t)kGCart)
    throw std::invalid_argument("a shell triple's Cartesian block exceeds the GPU kernel's LDS block");
}

}  // namespace

extern "C" {

int hfjk_int2c2e_device(hfjk_ctx* ctx, const hfjk_basis* aux, double* metric, void* stream) {
  return iguard([&] {
    if (ctx == nullptr) throw std::invalid_argument("null context");
    if (metric == nullptr) throw std::invalid_argument("null output pointer");
    int n = 0;
    const std::vector<Shell> sh = make_shells(aux, n);
    const std::vector<double> nm = norms(sh, n);
    check_item_sizes(sh, sh, true, nullptr);
    std::vector<DItem> items;
    for (int i = 0; i < (int)sh.size(); ++i)
      for (int j = 0; j <= i; ++j) items.push_back(DItem{i, -1, j});
    if (hipSetDevice(ctx->device) != hipSuccess) throw std::runtime_error("hipSetDevice failed");
    launch_triples(sh, sh, items, nm, nm, n, true, metric, (hipStream_t)stream);
  });
}

int hfjk_int3c2e_device(hfjk_ctx* ctx, const hfjk_basis* aux, const hfjk_basis* orbital, double* out, void* stream) {
  return iguard([&] {
    if (ctx == nullptr) throw std::invalid_argument("null context");
    if (out == nullptr) throw std::invalid_argument("null output pointer");
    int nx = 0, n = 0;
    const std::vector<Shell> sx = make_shells(aux, nx);
    const std::vector<Shell> so = make_shells(orbital, n);
    const std::vector<double> nmx = norms(sx, nx);
    const std::vector<double> nmo = norms(so, n);
    check_item_sizes(so, sx, false, &so);
    std::vector<DItem> items;
    for (int i = 0; i < (int)so.size(); ++i)
      for (int j = 0; j <= i; ++j) {
        if (!pair_survives(so[i], so[j])) continue;
        for (int c = 0; c < (int)sx.size(); ++c) items.push_back(DItem{i, j, c});
      }
    if (hipSetDevice(ctx->device) != hipSuccess) throw std::runtime_error("hipSetDevice failed");
    launch_triples(so, sx, items, nmo, nmx, n, false, out, (hipStream_t)stream);
  });
}

}  // extern "C"
