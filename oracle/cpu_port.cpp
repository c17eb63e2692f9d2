// C++/OpenMP CPU port of the reference ESP path -- TEST INFRASTRUCTURE ONLY
// (CPU baseline and large-N cross-check; never linked into the product).
//
//  cp_build : the table build of fastesp/native/build_pairs.cpp:162-425 (shell
//             pairs, folded density block, K cut, Hermite E, etab, G, bound,
//             reff, sum rule) with either screen: rule 0 = NumPy
//             (fastesp/table.py:190-194), rule 1 = C++ (build_pairs.cpp:318-328).
//             Pairs come out in reference order (per-shell-pair buffers are
//             concatenated in shell-pair order, not merged under a critical
//             section as in build_pairs.cpp:374-383).
//  cp_eval  : the per-point evaluation of fastesp/table.py:73-104 /
//             esp_gpu.cu:161-229 with the reference GPU Boys scheme
//             (esp_gpu.cu:36-66: Taylor series of F_M + downward recursion below
//             T = M - 1/2, erf + upward recursion above), OpenMP over points.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace {

struct LTab {
  int L = 0, T = 1;
  std::vector<std::array<int, 4>> keys;       // (ax, ay, az, m) sorted
  std::vector<std::vector<std::pair<int, double>>> rows;  // t (cube index) -> coef
  std::vector<std::array<int, 3>> monos;      // schedule
  std::vector<int> parent, axis, aidx;
  std::vector<double> wmom;
  std::vector<int> abs_t;
};

const double M_ABS[12] = {2.5066282746310002, 2.0, 1.5957691216057308, 4.0,
                          10.212920779565312, 33.63284766793385, 136.5963625280984,
                          665.1416330443616, 3768.863619586618, 25506.330595911374,
                          202331.81842358326, 1872154.2572862155};

LTab make_ltab(int L) {
  LTab lt;
  lt.L = L;
  lt.T = L + 1;
  const int T = lt.T;
  std::vector<std::vector<double>> a(L + 2, std::vector<double>(L / 2 + 2, 0.0));
  a[0][0] = 1.0;
  for (int t = 0; t < L + 1; ++t)
    for (int j = 0; j < L / 2 + 2; ++j) {
      double v = -a[t][j];
      if (j >= 1) v += (t - 2 * j + 2) * a[t][j - 1];
      a[t + 1][j] = v;
    }
  std::map<std::array<int, 4>, std::map<int, double>> acc;
  for (int tx = 0; tx < T; ++tx)
    for (int ty = 0; ty + tx < T; ++ty)
      for (int tz = 0; tz + ty + tx < T; ++tz) {
        const int lt_ = tx + ty + tz, col = (tx * T + ty) * T + tz;
        const double sgn = (lt_ & 1) ? -1.0 : 1.0;
        for (int jx = 0; jx <= tx / 2; ++jx)
          for (int jy = 0; jy <= ty / 2; ++jy)
            for (int jz = 0; jz <= tz / 2; ++jz) {
              const double c = a[tx][jx] * a[ty][jy] * a[tz][jz];
              if (c == 0.0) continue;
              acc[{tx - 2 * jx, ty - 2 * jy, tz - 2 * jz, lt_ - jx - jy - jz}][col] += sgn * c;
            }
      }
  for (auto& kv : acc) {
    lt.keys.push_back(kv.first);
    std::vector<std::pair<int, double>> r(kv.second.begin(), kv.second.end());
    lt.rows.push_back(r);
  }
  std::vector<std::array<int, 3>> grown;
  for (auto& k : lt.keys) {
    std::vector<std::array<int, 3>> st = {{k[0], k[1], k[2]}};
    while (!st.empty()) {
      auto u = st.back();
      st.pop_back();
      if (std::find(grown.begin(), grown.end(), u) != grown.end()) continue;
      grown.push_back(u);
      for (int d = 0; d < 3; ++d)
        if (u[d] > 0) {
          auto w = u;
          w[d] -= 1;
          st.push_back(w);
        }
    }
  }
  std::sort(grown.begin(), grown.end());
  lt.monos = grown;
  std::map<std::array<int, 3>, int> idx;
  for (size_t i = 0; i < grown.size(); ++i) idx[grown[i]] = (int)i;
  for (auto& m : grown) {
    int par = 0, ax = 0;
    for (int d = 0; d < 3; ++d)
      if (m[d] > 0) {
        auto w = m;
        w[d] -= 1;
        par = idx[w];
        ax = d;
        break;
      }
    lt.parent.push_back(par);
    lt.axis.push_back(ax);
  }
  for (auto& k : lt.keys) lt.aidx.push_back(idx[{k[0], k[1], k[2]}]);
  lt.wmom.resize(T * T * T);
  lt.abs_t.resize(T * T * T);
  for (int x = 0; x < T; ++x)
    for (int y = 0; y < T; ++y)
      for (int z = 0; z < T; ++z) {
        lt.wmom[(x * T + y) * T + z] = M_ABS[x] * M_ABS[y] * M_ABS[z];
        lt.abs_t[(x * T + y) * T + z] = x + y + z;
      }
  return lt;
}

const LTab& ltab(int L) {
  static std::vector<LTab> t = [] {
    std::vector<LTab> v;
    for (int L = 0; L <= 8; ++L) v.push_back(make_ltab(L));
    return v;
  }();
  return t[L];
}

struct Pair {
  double p, P[3], K;
  int L;
  int64_t goff;
};

struct Table {
  std::vector<Pair> pairs;
  std::vector<double> g;
  std::vector<double> Z, xyz;
  double sum_rule = 0.0;
  int64_t n_cand = 0;
};

// esp_gpu.cu:36-66 scheme
inline void boys_ref(double T, int M, double* F) {
  if (T < 1e-10) {
    for (int m = 0; m <= M; ++m) F[m] = 1.0 / (2 * m + 1);
    return;
  }
  if (T >= (double)M - 0.5) {
    const double F0 = (T > 30.0) ? 0.5 * std::sqrt(M_PI / T) : 0.5 * std::sqrt(M_PI / T) * std::erf(std::sqrt(T));
    F[0] = F0;
    const double e = std::exp(-T), h = 0.5 / T;
    for (int m = 0; m < M; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
  } else {
    double s = 0.0, tk = 1.0;
    for (int k = 0; k < 80; ++k) {
      const double add = tk / (2 * M + 2 * k + 1);
      s += add;
      if (std::fabs(add) <= 1e-18 * std::fabs(s)) break;
      tk *= -T / (k + 1);
    }
    F[M] = s;
    const double e = std::exp(-T);
    for (int m = M; m > 0; --m) F[m - 1] = (2.0 * T * F[m] + e) / (2 * m - 1);
  }
}

}  // namespace

extern "C" {

void* cp_build(int n_func, const int32_t* fn_atom, const int32_t* fn_nprim, const int32_t* fn_off,
               const int32_t* fn_lx, const int32_t* fn_ly, const int32_t* fn_lz,
               const int32_t* fn_shell, const double* fc, const double* pexp,
               const double* pcoef, const double* dens, int n_points, const double* pts,
               int n_atoms, const double* Z, const double* xyz, double tol, int rule,
               int nthreads) {
  struct Sh { int l, fn0, nfn, prim0, nprim; double c[3]; };
  std::vector<Sh> sh;
  for (int i = 0; i < n_func;) {
    int j = i;
    while (j < n_func && fn_shell[j] == fn_shell[i]) ++j;
    Sh s{fn_lx[i] + fn_ly[i] + fn_lz[i], i, j - i, fn_off[i], fn_nprim[i],
         {fc[3 * i], fc[3 * i + 1], fc[3 * i + 2]}};
    sh.push_back(s);
    i = j;
  }
  const int nsh = (int)sh.size();
  std::vector<std::pair<int, int>> sps;
  for (int i = 0; i < nsh; ++i)
    for (int j = i; j < nsh; ++j) sps.push_back({i, j});
  const int nsp = (int)sps.size();
  std::vector<std::vector<Pair>> out_p(nsp);
  std::vector<std::vector<double>> out_g(nsp);
  std::vector<double> srule(nsp, 0.0);
  std::vector<int64_t> ncand(nsp, 0);
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int s = 0; s < nsp; ++s) {
    const Sh& A = sh[sps[s].first];
    const Sh& B = sh[sps[s].second];
    const bool same = sps[s].first == sps[s].second;
    std::vector<double> blk(A.nfn * B.nfn);
    double mx = 0.0;
    for (int mi = 0; mi < A.nfn; ++mi)
      for (int mj = 0; mj < B.nfn; ++mj) {
        double d = dens[(size_t)(A.fn0 + mi) * n_func + (B.fn0 + mj)];
        if (!same) d += dens[(size_t)(B.fn0 + mj) * n_func + (A.fn0 + mi)];
        blk[mi * B.nfn + mj] = d;
        mx = std::max(mx, std::fabs(d));
      }
    if (mx < 1e-15) continue;
    const int L = A.l + B.l, T = L + 1;
    const LTab& lt = ltab(L);
    const int nterm = (int)lt.keys.size();
    const double dx = A.c[0] - B.c[0], dy = A.c[1] - B.c[1], dz = A.c[2] - B.c[2];
    const double ab2 = dx * dx + dy * dy + dz * dz;
    std::vector<double> etab(T * T * T), E[3];
    const int d1 = B.l + 1;
    for (int ii = 0; ii < A.nprim; ++ii)
      for (int jj = 0; jj < B.nprim; ++jj) {
        const double al = pexp[A.prim0 + ii], be = pexp[B.prim0 + jj], p = al + be;
        const double cc = pcoef[A.prim0 + ii] * pcoef[B.prim0 + jj];
        const double K = std::exp(-al * be / p * ab2);
        if (K < 1e-30) continue;
        ncand[s]++;
        double P[3], xa[3], xb[3];
        for (int k = 0; k < 3; ++k) {
          P[k] = (al * A.c[k] + be * B.c[k]) / p;
          xa[k] = P[k] - A.c[k];
          xb[k] = P[k] - B.c[k];
        }
        for (int ax = 0; ax < 3; ++ax) {
          auto& e = E[ax];
          e.assign((A.l + 1) * d1 * T, 0.0);
          e[0] = 1.0;
          for (int i = 0; i <= A.l; ++i)
            for (int j = 0; j <= B.l; ++j) {
              if (!i && !j) continue;
              const double* src = i ? &e[((i - 1) * d1 + j) * T] : &e[(i * d1 + j - 1) * T];
              const double x = i ? xa[ax] : xb[ax];
              double* dst = &e[(i * d1 + j) * T];
              for (int t = 0; t < T; ++t) {
                double v = x * src[t];
                if (t >= 1) v += src[t - 1] / (2.0 * p);
                if (t + 1 < T) v += (t + 1) * src[t + 1];
                dst[t] = v;
              }
            }
        }
        std::fill(etab.begin(), etab.end(), 0.0);
        for (int mi = 0; mi < A.nfn; ++mi)
          for (int mj = 0; mj < B.nfn; ++mj) {
            const double w = blk[mi * B.nfn + mj] * cc;
            if (w == 0.0) continue;
            const int f = A.fn0 + mi, g = B.fn0 + mj;
            const double* ex = &E[0][(fn_lx[f] * d1 + fn_lx[g]) * T];
            const double* ey = &E[1][(fn_ly[f] * d1 + fn_ly[g]) * T];
            const double* ez = &E[2][(fn_lz[f] * d1 + fn_lz[g]) * T];
            for (int tx = 0; tx < T; ++tx)
              for (int ty = 0; ty < T; ++ty) {
                const double cxy = w * ex[tx] * ey[ty];
                if (cxy == 0.0) continue;
                for (int tz = 0; tz < T; ++tz) etab[(tx * T + ty) * T + tz] += cxy * ez[tz];
              }
          }
        srule[s] += K * std::pow(M_PI / p, 1.5) * etab[0];
        std::vector<double> G(nterm);
        bool anyg = false;
        for (int k = 0; k < nterm; ++k) {
          double v = 0.0;
          for (auto& rc : lt.rows[k]) v += rc.second * etab[rc.first];
          G[k] = 2.0 * M_PI / p * std::pow(2.0 * p, lt.keys[k][3]) * v;
          anyg |= (G[k] != 0.0);
        }
        double bound = 0.0;
        for (int c = 0; c < T * T * T; ++c)
          bound += std::fabs(etab[c]) * lt.wmom[c] * std::pow(2.0 * p, 0.5 * lt.abs_t[c]);
        bound *= std::pow(2.0 * p, -1.5);
        const double reff = 1.6 * std::sqrt((L + 2.0) / p);
        double d2 = INFINITY;
        for (int q = 0; q < n_points; ++q) {
          const double x = pts[3 * q] - P[0], y = pts[3 * q + 1] - P[1], z = pts[3 * q + 2] - P[2];
          d2 = std::min(d2, x * x + y * y + z * z);
        }
        double dd = std::sqrt(d2);
        if (rule == 1) {
          if (!anyg || cc == 0.0) continue;
          dd -= reff;
        }
        if (dd < 0.05) dd = 0.05;
        if (K * bound / dd < tol) continue;
        Pair pr{p, {P[0], P[1], P[2]}, K, L, 0};
        pr.goff = (int64_t)out_g[s].size();
        out_p[s].push_back(pr);
        out_g[s].insert(out_g[s].end(), G.begin(), G.end());
      }
  }
  Table* t = new Table();
  for (int s = 0; s < nsp; ++s) {
    const int64_t base = (int64_t)t->g.size();
    for (auto pr : out_p[s]) {
      pr.goff += base;
      t->pairs.push_back(pr);
    }
    t->g.insert(t->g.end(), out_g[s].begin(), out_g[s].end());
    t->sum_rule += srule[s];
    t->n_cand += ncand[s];
  }
  t->Z.assign(Z, Z + n_atoms);
  t->xyz.assign(xyz, xyz + 3 * n_atoms);
  return t;
}

int cp_info(void* h, int64_t* out) {
  Table* t = (Table*)h;
  out[0] = (int64_t)t->pairs.size();
  out[1] = (int64_t)t->g.size();
  out[2] = t->n_cand;
  return 0;
}

double cp_sum_rule(void* h) { return ((Table*)h)->sum_rule; }

// per-pair p, Px, Py, Pz, K, L  (n_pairs x 6) and the flat g array
int cp_export(void* h, double* pairs6, double* g) {
  Table* t = (Table*)h;
  for (size_t i = 0; i < t->pairs.size(); ++i) {
    const Pair& p = t->pairs[i];
    double* r = pairs6 + 6 * i;
    r[0] = p.p; r[1] = p.P[0]; r[2] = p.P[1]; r[3] = p.P[2]; r[4] = p.K; r[5] = p.L;
  }
  std::copy(t->g.begin(), t->g.end(), g);
  return 0;
}

int cp_eval(void* h, const double* pts, int n, double* V, int nthreads) {
  Table* t = (Table*)h;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int np = (int)t->pairs.size();
  const int nat = (int)t->Z.size();
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < n; ++i) {
    const double cx = pts[3 * i], cy = pts[3 * i + 1], cz = pts[3 * i + 2];
    double acc = 0.0;
    double F[16], mono[192];
    for (int ip = 0; ip < np; ++ip) {
      const Pair& pr = t->pairs[ip];
      const LTab& lt = ltab(pr.L);
      const double R[3] = {cx - pr.P[0], cy - pr.P[1], cz - pr.P[2]};
      const double T = pr.p * (R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
      boys_ref(T, pr.L, F);
      mono[0] = 1.0;
      const int nm = (int)lt.monos.size();
      for (int j = 1; j < nm; ++j) mono[j] = mono[lt.parent[j]] * R[lt.axis[j]];
      const double* g = &t->g[pr.goff];
      double s = 0.0;
      const int nt = (int)lt.keys.size();
      for (int k = 0; k < nt; ++k) s += g[k] * mono[lt.aidx[k]] * F[lt.keys[k][3]];
      acc += pr.K * s;
    }
    double nuc = 0.0;
    for (int a = 0; a < nat; ++a) {
      const double dx = cx - t->xyz[3 * a], dy = cy - t->xyz[3 * a + 1], dz = cz - t->xyz[3 * a + 2];
      nuc += t->Z[a] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    V[i] = nuc - acc;
  }
  return 0;
}

void cp_free(void* h) { delete (Table*)h; }

}  // extern "C"
