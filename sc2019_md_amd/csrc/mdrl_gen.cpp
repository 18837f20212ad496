// mdrl_gen.cpp -- seeded synthetic systems of the benchmark workloads.
//
// One generator serves the GPU runs and every CPU baseline so that both see identical
// inputs.  Random numbers: SplitMix64 seeds xoshiro256**; uniforms are the top 53 bits;
// normals are Box-Muller pairs.  Lattice site (i, j, k) sits at ((i, j, k) + 0.5) * a +
// U(-jitter, +jitter)^3.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mdrl_internal.h"

using mdrl::set_error;

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (int k = 0; k < 4; ++k) s[k] = splitmix(x);
    }
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal()
    {
        if (has_spare) { has_spare = false; return spare; }
        double u1 = 1.0 - uniform(), u2 = uniform();
        double r = sqrt(-2.0 * log(u1)), th = 6.283185307179586 * u2;
        spare = r * sin(th);
        has_spare = true;
        return r * cos(th);
    }
    uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n); }
    bool has_spare = false;
    double spare = 0.0;
};

constexpr double kB = 0.0083145;          // kJ/mol/K
constexpr double kCoulomb = 1389.35;      // kJ A / (mol e^2)

// Mixed-species recipe: types with fractions 0.50/0.25/0.15/0.10.
const float kSig4[4] = {1.0f, 3.2f, 3.5f, 3.0f};
const float kEps4[4] = {0.19f, 0.65f, 0.30f, 0.71f};
// Masses are not part of the workload definition; every type gets the argon mass of the
// 512-atom argon workload.  The stiffest contacts are Coulomb-bound pairs of the sigma 1 A
// type (charges drawn N(0, 0.4^2) e reach ~1.8 e among 92,224 atoms, ~2 e among 1,048,576),
// and velocity Verlet needs omega*dt < 2: with carbon-like masses (12-16 amu) the 92,224-
// and 1,048,576-atom systems diverge within 50 steps at dt = 1 fs, with 39.948 amu they hold
// their total energy to <1% over 1,200 steps (the mass scales omega by 1/sqrt(m)).
const float kMass4[4] = {39.948f, 39.948f, 39.948f, 39.948f};

void shuffle(std::vector<int> &v, Rng &r)
{
    for (size_t i = v.size(); i > 1; --i) {
        size_t j = (size_t)r.below(i);
        std::swap(v[i - 1], v[j]);
    }
}

// n atoms on a jittered nx x ny x nz lattice of the box, sites chosen by permutation.
void lattice(md_rl_system *s, int nx, int ny, int nz, double jitter, Rng &r)
{
    const int nsite = nx * ny * nz;
    std::vector<int> sites(nsite);
    for (int k = 0; k < nsite; ++k) sites[k] = k;
    if (s->n < nsite) {
        shuffle(sites, r);
        sites.resize(s->n);
        std::sort(sites.begin(), sites.end());
    }
    const double ax = s->box[0] / nx, ay = s->box[1] / ny, az = s->box[2] / nz;
    for (int a = 0; a < s->n; ++a) {
        int k = sites[a];
        int i = k / (ny * nz), j = (k / nz) % ny, l = k % nz;
        double p[3] = {(i + 0.5) * ax, (j + 0.5) * ay, (l + 0.5) * az};
        for (int d = 0; d < 3; ++d) {
            double x = p[d] + (2.0 * r.uniform() - 1.0) * jitter;
            x -= s->box[d] * floor(x / s->box[d]);
            float f = (float)x;
            if (f >= s->box[d]) f = 0.0f;
            s->particle[11 * a + d] = f;
        }
    }
}

void maxwell(md_rl_system *s, double T, Rng &r)
{
    double p[3] = {0, 0, 0}, mt = 0;
    std::vector<double> v(3 * (size_t)s->n);
    for (int a = 0; a < s->n; ++a) {
        double m = s->mass[s->type[a]];
        double sd = sqrt(kB * T / m) * 0.01;   // sqrt(kJ/mol/amu) = 1 nm/ps = 0.01 A/fs
        for (int d = 0; d < 3; ++d) { v[3 * a + d] = r.normal() * sd; p[d] += m * v[3 * a + d]; }
        mt += m;
    }
    for (int a = 0; a < s->n; ++a)
        for (int d = 0; d < 3; ++d) s->particle[11 * a + 3 + d] = (float)(v[3 * a + d] - p[d] / mt);
}

void species4(md_rl_system *s, Rng &r)
{
    s->ntypes = 4;
    for (int t = 0; t < 4; ++t) { s->sigma[t] = kSig4[t]; s->eps[t] = kEps4[t]; s->mass[t] = kMass4[t]; }
    const int n = s->n;
    int cnt[4];
    cnt[1] = (int)lround(0.25 * n);
    cnt[2] = (int)lround(0.15 * n);
    cnt[3] = (int)lround(0.10 * n);
    cnt[0] = n - cnt[1] - cnt[2] - cnt[3];
    std::vector<int> ty;
    ty.reserve(n);
    for (int t = 0; t < 4; ++t) ty.insert(ty.end(), cnt[t], t);
    shuffle(ty, r);
    for (int a = 0; a < n; ++a) s->type[a] = ty[a];
    double mean = 0;
    std::vector<double> q(n);
    for (int a = 0; a < n; ++a) { q[a] = 0.4 * r.normal(); mean += q[a]; }
    mean /= n;
    for (int a = 0; a < n; ++a) s->charge[a] = (float)(q[a] - mean);
    s->coulomb_k = (float)kCoulomb;
}

int alloc(md_rl_system *s, int n)
{
    memset(s, 0, sizeof *s);
    s->n = n;
    s->particle = (float *)calloc((size_t)n * 11, sizeof(float));
    s->type = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    s->charge = (float *)calloc((size_t)n, sizeof(float));
    if (!s->particle || !s->type || !s->charge) {
        free(s->particle); free(s->type); free(s->charge);
        set_error("md_rl_make_system: out of host memory for %d atoms", n);
        return MDRL_ENOMEM;
    }
    return MDRL_OK;
}

// Size of replica `index` of the batched set: log-uniform in [256, 16384].
int replica_size(int index, uint64_t seed)
{
    Rng r(seed ^ 0x5bd1e995ull);
    int n = 0;
    for (int k = 0; k <= index; ++k) n = (int)lround(exp(log(256.0) + r.uniform() * (log(16384.0) - log(256.0))));
    return n;
}

// Workload 5: 64 linear chains of 48 united-atom beads at 0.033 beads/A^3, grown bead by
// bead (bond 1.53 A, angle 111 deg, torsion trans 60% / gauche +-65 deg 20% each, with
// 8 deg of noise) while keeping every bead >= 2.5 A from all beads it is not excluded
// from; a bead that finds no free torsion in 30 tries restarts its chain.
constexpr int kChains5 = 64, kBeads5 = 48;
constexpr double kBond5 = 1.53, kTheta5 = 111.0 * 3.14159265358979323846 / 180.0;

struct Grid {
    double L, cs;
    int m;
    std::vector<std::vector<int>> cells;
    Grid(double L_, double dmin) : L(L_)
    {
        m = std::max(1, (int)floor(L / dmin));
        cs = L / m;
        cells.resize((size_t)m * m * m);
    }
    int cell(const double *x) const
    {
        int c[3];
        for (int d = 0; d < 3; ++d) c[d] = std::min(m - 1, std::max(0, (int)(x[d] / cs)));
        return (c[0] * m + c[1]) * m + c[2];
    }
};

double wrap_d(double x, double L) { return x - L * floor(x / L); }

// True when x keeps dmin from every placed bead except the previous `skip` beads of its
// own chain (indices first..self-1 within skip).
bool free_at(const Grid &g, const std::vector<double> &pos, const double *x, int self, int chain_first,
             double dmin)
{
    int c[3];
    for (int d = 0; d < 3; ++d) c[d] = std::min(g.m - 1, std::max(0, (int)(x[d] / g.cs)));
    const int span = g.m >= 3 ? 1 : 0;
    for (int ox = -span; ox <= span; ++ox)
        for (int oy = -span; oy <= span; ++oy)
            for (int oz = -span; oz <= span; ++oz) {
                const int cx = (c[0] + ox + g.m) % g.m, cy = (c[1] + oy + g.m) % g.m, cz = (c[2] + oz + g.m) % g.m;
                for (int b : g.cells[(size_t)(cx * g.m + cy) * g.m + cz]) {
                    if (b >= chain_first && self - b <= 3) continue;
                    double r2 = 0;
                    for (int d = 0; d < 3; ++d) {
                        double dx = x[d] - pos[3 * (size_t)b + d];
                        dx -= g.L * floor(dx / g.L + 0.5);
                        r2 += dx * dx;
                    }
                    if (r2 < dmin * dmin) return false;
                }
            }
    return true;
}

void unit(double *v)
{
    const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int d = 0; d < 3; ++d) v[d] /= l;
}
void crossd(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

int grow_chains(md_rl_system *s, Rng &r)
{
    const double L = s->box[0];
    double dmin = 2.5;
    Grid g(L, dmin);
    std::vector<double> pos(3 * (size_t)s->n, 0.0);   // unwrapped along each chain
    std::vector<double> wp(3 * (size_t)s->n, 0.0);    // wrapped
    for (int ch = 0; ch < kChains5; ++ch) {
        const int first = ch * kBeads5;
        int restarts = 0;
        for (;;) {
            bool ok = true;
            for (int k = 0; k < kBeads5 && ok; ++k) {
                const int i = first + k;
                bool placed = false;
                for (int attempt = 0; attempt < 30 && !placed; ++attempt) {
                    double x[3];
                    if (k == 0) {
                        for (int d = 0; d < 3; ++d) x[d] = r.uniform() * L;
                    } else if (k == 1) {
                        double u[3] = {r.normal(), r.normal(), r.normal()};
                        unit(u);
                        for (int d = 0; d < 3; ++d) x[d] = pos[3 * (size_t)(i - 1) + d] + kBond5 * u[d];
                    } else {
                        const double *B = &pos[3 * (size_t)(i - 2)], *C = &pos[3 * (size_t)(i - 1)];
                        double bc[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
                        unit(bc);
                        double ref[3];
                        if (k >= 3) {
                            const double *A = &pos[3 * (size_t)(i - 3)];
                            for (int d = 0; d < 3; ++d) ref[d] = B[d] - A[d];
                        } else {
                            for (int d = 0; d < 3; ++d) ref[d] = r.normal();
                        }
                        double nv[3];
                        crossd(ref, bc, nv);
                        unit(nv);
                        double mv[3];
                        crossd(nv, bc, mv);
                        double phi;
                        const double u = r.uniform();
                        if (attempt == 0 && k >= 3)
                            phi = (u < 0.6 ? 180.0 : u < 0.8 ? 65.0 : -65.0) + 8.0 * r.normal();
                        else
                            phi = 360.0 * u;
                        phi *= 3.14159265358979323846 / 180.0;
                        const double a = -kBond5 * cos(kTheta5), b = kBond5 * sin(kTheta5) * cos(phi),
                                     c = kBond5 * sin(kTheta5) * sin(phi);
                        for (int d = 0; d < 3; ++d) x[d] = C[d] + a * bc[d] + b * mv[d] + c * nv[d];
                    }
                    double w[3];
                    for (int d = 0; d < 3; ++d) w[d] = wrap_d(x[d], L);
                    if (free_at(g, wp, w, i, first, dmin)) {
                        for (int d = 0; d < 3; ++d) { pos[3 * (size_t)i + d] = x[d]; wp[3 * (size_t)i + d] = w[d]; }
                        g.cells[g.cell(w)].push_back(i);
                        placed = true;
                    }
                }
                if (!placed) ok = false;
            }
            if (ok) break;
            // take the chain's beads out of the grid and start it again
            for (auto &cell : g.cells)
                cell.erase(std::remove_if(cell.begin(), cell.end(), [&](int b) { return b >= first && b < first + kBeads5; }),
                           cell.end());
            if (++restarts % 200 == 0) dmin *= 0.95;
            if (restarts > 5000) { set_error("md_rl_make_system: chain growth failed"); return MDRL_EINVAL; }
        }
    }
    for (int i = 0; i < s->n; ++i)
        for (int d = 0; d < 3; ++d) {
            float f = (float)wp[3 * (size_t)i + d];
            if (f >= s->box[d]) f = 0.0f;
            s->particle[11 * (size_t)i + d] = f;
        }
    return MDRL_OK;
}

int chain_topology(md_rl_system *s)
{
    const int per = (kBeads5 - 1) + (kBeads5 - 2) + (kBeads5 - 3);
    s->nterms = kChains5 * per;
    s->term_atoms = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)s->nterms);
    s->term_params = (float *)malloc(sizeof(float) * 4 * (size_t)s->nterms);
    s->excl = (uint32_t *)calloc((size_t)s->n, sizeof(uint32_t));
    if (!s->term_atoms || !s->term_params || !s->excl) { set_error("md_rl_make_system: out of host memory"); return MDRL_ENOMEM; }
    int t = 0;
    auto put = [&](int a, int b, int c, int d, float p0, float p1, float p2) {
        int32_t *q = s->term_atoms + 4 * (size_t)t;
        float *w = s->term_params + 4 * (size_t)t;
        q[0] = a; q[1] = b; q[2] = c; q[3] = d;
        w[0] = p0; w[1] = p1; w[2] = p2; w[3] = 0.0f;
        ++t;
    };
    // U = k (r - r0)^2 with k = 2.2e5 / 2 kJ/mol/nm^2 = 1100 kJ/mol/A^2 (a harmonic bond of
    // force constant 2.2e5 kJ/mol/nm^2); angles 520 / 2 kJ/mol/rad^2 likewise.
    for (int ch = 0; ch < kChains5; ++ch) {
        const int f = ch * kBeads5;
        for (int k = 0; k + 1 < kBeads5; ++k) put(f + k, f + k + 1, -1, -1, 1.53f, 1100.0f, 0.0f);
        for (int k = 0; k + 2 < kBeads5; ++k) put(f + k, f + k + 1, f + k + 2, -1, (float)kTheta5, 260.0f, 0.0f);
        for (int k = 0; k + 3 < kBeads5; ++k) put(f + k, f + k + 1, f + k + 2, f + k + 3, 5.9f, 3.0f, 0.0f);
        for (int k = 0; k < kBeads5; ++k)
            for (int dk = -3; dk <= 3; ++dk)
                if (dk != 0 && k + dk >= 0 && k + dk < kBeads5) s->excl[f + k] |= 1u << (dk + 16);
    }
    return MDRL_OK;
}

}  // namespace

extern "C" int md_rl_make_system(md_rl_system *s, int which, int index, uint64_t seed)
{
    if (!s) { set_error("md_rl_make_system: NULL system"); return MDRL_EINVAL; }
    int rc;
    switch (which) {
    case 1: {   // 512 argon atoms, 2x2x2 per cell of a 4^3 grid of 8.5 A cells
        if ((rc = alloc(s, 512)) != MDRL_OK) return rc;
        Rng r(seed);
        s->box[0] = s->box[1] = s->box[2] = 34.0f;
        s->rc = 8.5f; s->dt = 2.0f; s->ntypes = 1; s->coulomb_k = 0.0f;
        s->sigma[0] = 3.4f; s->eps[0] = 0.996f; s->mass[0] = 39.948f;
        lattice(s, 8, 8, 8, 0.3, r);
        maxwell(s, 94.0, r);
        return MDRL_OK;
    }
    case 2: {   // 92,224-atom LJ + Coulomb, 108.86 x 108.86 x 77.76 A
        if ((rc = alloc(s, 92224)) != MDRL_OK) return rc;
        Rng r(seed);
        s->box[0] = 108.86f; s->box[1] = 108.86f; s->box[2] = 77.76f;
        s->rc = 12.0f; s->dt = 1.0f;
        lattice(s, 51, 51, 36, 0.4, r);
        species4(s, r);
        maxwell(s, 300.0, r);
        return MDRL_OK;
    }
    case 3: {   // 1,048,576 atoms in a 219.0 A cube
        if ((rc = alloc(s, 1048576)) != MDRL_OK) return rc;
        Rng r(seed);
        s->box[0] = s->box[1] = s->box[2] = 219.0f;
        s->rc = 12.0f; s->dt = 1.0f;
        lattice(s, 102, 102, 102, 0.4, r);
        species4(s, r);
        maxwell(s, 300.0, r);
        return MDRL_OK;
    }
    case 4: {   // replica `index` of the 512-system batch: 0.1 atoms/A^3, rc = 10 A
        if (index < 0) { set_error("md_rl_make_system: replica index < 0"); return MDRL_EINVAL; }
        const int n = replica_size(index, seed);
        if ((rc = alloc(s, n)) != MDRL_OK) return rc;
        Rng r(Rng::splitmix(seed += 0x632be59bd9b4e019ull * (uint64_t)(index + 1)));
        const float L = (float)cbrt(n / 0.1);
        s->box[0] = s->box[1] = s->box[2] = L;
        s->rc = 10.0f; s->dt = 1.0f;
        const int m = (int)ceil(cbrt((double)n) - 1e-9);
        lattice(s, m, m, m, 0.4, r);
        species4(s, r);
        maxwell(s, 300.0, r);
        return MDRL_OK;
    }
    case 5: {   // replica `index` of the 128 bonded systems: 64 chains x 48 beads
        if (index < 0) { set_error("md_rl_make_system: replica index < 0"); return MDRL_EINVAL; }
        const int n = kChains5 * kBeads5;
        if ((rc = alloc(s, n)) != MDRL_OK) return rc;
        Rng r(Rng::splitmix(seed += 0x2545f4914f6cdd1dull * (uint64_t)(index + 1)));
        const float L = (float)cbrt(n / 0.033);
        s->box[0] = s->box[1] = s->box[2] = L;
        s->rc = 10.0f; s->dt = 2.0f;
        s->ntypes = 1;
        s->sigma[0] = 3.9f; s->eps[0] = 0.48f; s->mass[0] = 14.027f;
        s->coulomb_k = (float)kCoulomb;
        if ((rc = grow_chains(s, r)) != MDRL_OK || (rc = chain_topology(s)) != MDRL_OK) {
            md_rl_free_system(s);
            return rc;
        }
        for (int i = 0; i < n; ++i) s->charge[i] = ((i % kBeads5) % 2 == 0) ? 0.25f : -0.25f;
        maxwell(s, 300.0, r);
        return MDRL_OK;
    }
    default:
        set_error("md_rl_make_system: unknown workload %d", which);
        return MDRL_EINVAL;
    }
}

extern "C" void md_rl_free_system(md_rl_system *s)
{
    if (!s) return;
    free(s->particle); free(s->type); free(s->charge);
    free(s->term_atoms); free(s->term_params); free(s->excl);
    s->particle = nullptr; s->type = nullptr; s->charge = nullptr;
    s->term_atoms = nullptr; s->term_params = nullptr; s->excl = nullptr;
    s->nterms = 0;
}
