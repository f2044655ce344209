// kp_mesh.cpp -- host block tree, ids, geometry and face-neighbour table.
//
// Behaviour follows the reference mesh layer (paths relative to src/):
//   tree_t::insert / neighbors ........ mesh/tree/tree.cpp:89-218
//   region_logic_t::coarsen ........... mesh/geometry/region.h:61-67
//   neighbor_t (offset, mode, d_lvl) .. mesh/geometry/region.h:110-146
//   lexicographic block order ......... mesh/geometry/region.h:17-33
//   dist_default ids .................. mesh/meshgen/meshgen.cpp:168-188
//   block axes / volume ............... mesh/meshgen/meshgen.cpp:252-351
//   physical-boundary flags ........... mesh/meshgen/meshgen.cpp:390-398
//   SMR region -> block index ......... mesh/meshgen/meshgen.cpp:111-142,353-373
//   face table (proxy setup) .......... modules/particle/types/proxy.h:22-62
//   spherical axes / volumes .......... mesh/meshgen/spherical.h:9-140,
//                                       mesh/meshgen/cylindrical.h:18-43
// The tree is a std::map keyed in the same (level, x0, x1, x2) order, so the
// ids handed out by kp_mesh_finalize are the reference's dist_mode=0 ids.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "kp_error.h"

namespace {

struct reg_t {
    int level = 0;
    std::array<int, 3> x{0, 0, 0};
};

struct lex_less {
    bool operator()(const reg_t &l, const reg_t &r) const
    {
        if (l.level != r.level)
            return l.level < r.level;
        for (int i = 0; i < 3; ++i)
            if (l.x[i] != r.x[i])
                return l.x[i] < r.x[i];
        return false;
    }
};

struct node_t {
    int id = -1;
    bool mask = false;
};

// floor division by 2 (region_logic_t::coarsen for negative coordinates)
inline int half_floor(int v) { return (v > 0 ? v : v - 1) / 2; }

reg_t coarser(reg_t r)
{
    r.level -= 1;
    for (auto &v : r.x)
        v = half_floor(v);
    return r;
}

// neighbor_t of region.h:110-140: offset of guest g seen from host h in the
// refined frame of the finer of the two.
struct nbr_t {
    std::array<int, 3> x{0, 0, 0};
    int d_lvl = 0;
    int mode = -1;
};

int shift_floor(int v, int s) // v >> s with floor semantics
{
    while (s-- > 0)
        v = (v >= 0) ? (v / 2) : -((-v + 1) / 2);
    return v;
}

nbr_t make_nbr(const reg_t &h, const reg_t &g)
{
    nbr_t nb;
    nb.d_lvl = g.level - h.level;
    for (int n = 0; n < 3; ++n) {
        int base = nb.d_lvl >= 0 ? (h.x[n] * (1 << nb.d_lvl))
                                 : shift_floor(h.x[n], -nb.d_lvl);
        int v = g.x[n] - base;
        if (v > 0 && nb.d_lvl <= 0)
            v *= 2;
        nb.x[n] = v;
    }
    nb.mode = -1;
    for (int n = 0; n < 3; ++n)
        nb.mode += (nb.x[n] != 0 && nb.x[n] != 1) ? 1 : 0;
    return nb;
}

} // namespace

struct kp_mesh {
    std::array<int, 3> reg_max{1, 1, 1};
    int n_cell = 16;
    std::array<int, 3> periodic{0, 0, 0};
    std::array<double, 3> x_min{0, 0, 0}, x_max{1, 1, 1};
    std::map<reg_t, node_t, lex_less> tree;
    bool finalized = false;
    int coord = KP_COORD_CARTESIAN;
    double ratio = 1.0; // r width ratio of consecutive root-level cells
    double dr_0 = -1.0; // first root-level r width (> 0: non-uniform r)

    // block table
    std::vector<reg_t> blocks;
    std::vector<double> xf0, dx0, vol0, xf, sph;
    std::vector<int32_t> level, logic, nb_id;
    std::vector<uint8_t> nb_shift, phybnd;
    std::vector<int8_t> d_lvl;

    std::array<int, 3> region_max(int lvl) const
    {
        return {reg_max[0] << lvl, reg_max[1] << lvl, reg_max[2] << lvl};
    }
    bool is_inside(const reg_t &r) const
    {
        auto rl = region_max(r.level);
        for (int n = 0; n < 3; ++n)
            if (r.x[n] < 0 || r.x[n] >= rl[n])
                return false;
        return true;
    }
    void regularize(reg_t &r) const
    {
        auto rl = region_max(r.level);
        for (int n = 0; n < 3; ++n)
            if (periodic[n])
                r.x[n] = ((r.x[n] % rl[n]) + rl[n]) % rl[n];
    }
    bool contains(const reg_t &r) const { return tree.count(r) != 0; }

    // tree.cpp:202-218
    void insert(const reg_t &reg)
    {
        if (contains(reg))
            return;
        if (reg.level > 0) {
            const reg_t rc = coarser(reg);
            insert(rc);
            tree[rc].mask = true;
            for (int d0 = 0; d0 < 2; ++d0)
                for (int d1 = 0; d1 < 2; ++d1)
                    for (int d2 = 0; d2 < 2; ++d2) {
                        reg_t ch;
                        ch.level = rc.level + 1;
                        ch.x = {2 * rc.x[0] + d0, 2 * rc.x[1] + d1, 2 * rc.x[2] + d2};
                        tree[ch];
                    }
            for (int d0 = -1; d0 <= 1; ++d0)
                for (int d1 = -1; d1 <= 1; ++d1)
                    for (int d2 = -1; d2 <= 1; ++d2) {
                        if (!d0 && !d1 && !d2)
                            continue;
                        reg_t t = rc;
                        t.x = {rc.x[0] + d0, rc.x[1] + d1, rc.x[2] + d2};
                        regularize(t);
                        if (is_inside(t))
                            insert(t);
                    }
        }
        tree[reg];
    }

    // tree.cpp:177-197: for every refined probe around reg, walk up to the
    // first existing unmasked node (not coarser than reg.level-1).
    // Returns (neighbour offset, guest region unregularized, guest leaf).
    struct found_t {
        nbr_t nb;
        reg_t guest;     // leaf as stored in the tree (regularized)
        reg_t guest_raw; // coarsened probe, not regularized (periodic images)
    };
    std::vector<found_t> neighbors(const reg_t &reg) const
    {
        std::map<std::array<int, 4>, found_t> res; // keyed like lex_map_t<neighbor_t>
        const int l_min = std::max(0, reg.level - 1);
        for (int d0 = -1; d0 <= 2; ++d0)
            for (int d1 = -1; d1 <= 2; ++d1)
                for (int d2 = -1; d2 <= 2; ++d2) {
                    const bool inside_blk = (d0 >= 0 && d0 <= 1) && (d1 >= 0 && d1 <= 1) &&
                                            (d2 >= 0 && d2 <= 1);
                    if (inside_blk)
                        continue;
                    reg_t tst;
                    tst.level = reg.level + 1;
                    tst.x = {d0 + 2 * reg.x[0], d1 + 2 * reg.x[1], d2 + 2 * reg.x[2]};
                    const reg_t orig = tst;
                    regularize(tst);
                    if (!is_inside(tst))
                        continue;
                    reg_t t = tst, t0 = orig;
                    for (; t.level >= l_min; t = coarser(t), t0 = coarser(t0)) {
                        auto it = tree.find(t);
                        if (it != tree.end() && !it->second.mask) {
                            found_t f;
                            f.nb = make_nbr(reg, t0);
                            f.guest_raw = t0;
                            f.guest = t0;
                            regularize(f.guest);
                            res[{0, f.nb.x[0], f.nb.x[1], f.nb.x[2]}] = f;
                            break;
                        }
                    }
                }
        std::vector<found_t> out;
        for (auto &p : res)
            out.push_back(p.second);
        return out;
    }

    // meshgen::base_t::location (meshgen.cpp:252-258); the non-uniform r axis
    // of cyl_t::location (cylindrical.h:34-43): r_min + dr_0 (q^(n x) - 1)/(q - 1)
    double location(double x_logic, int axis) const
    {
        if (axis != 0 || !(dr_0 > 0.0))
            return x_min[axis] + x_logic * (x_max[axis] - x_min[axis]);
        const double n_logic = x_logic * (double)(reg_max[0] * n_cell);
        return x_min[0] + dr_0 / (ratio - 1.0) * (std::pow(ratio, n_logic) - 1.0);
    }

    // meshgen.cpp:353-373
    int search_blk_idx(double loc, int lvl, int axis) const
    {
        if (loc > x_max[axis] || loc < x_min[axis])
            return -1;
        const int i_tot = reg_max[axis] << lvl;
        int i_r = i_tot, i_l = 0;
        const double i_ftot = (double)i_tot;
        while (i_r - i_l > 1) {
            const int i_m = (i_l + i_r) / 2;
            const double xl = location(i_m / i_ftot, axis);
            if (xl < loc)
                i_l = i_m;
            else
                i_r = i_m;
        }
        return i_l;
    }

    void finalize()
    {
        // fill_base_level (tree.cpp:220-230)
        for (int i2 = 0; i2 < reg_max[2]; ++i2)
            for (int i1 = 0; i1 < reg_max[1]; ++i1)
                for (int i0 = 0; i0 < reg_max[0]; ++i0) {
                    reg_t r;
                    r.level = 0;
                    r.x = {i0, i1, i2};
                    tree[r];
                }
        // dist_default (meshgen.cpp:168-188), single rank
        blocks.clear();
        int id = 0;
        for (auto &p : tree) {
            if (p.second.mask)
                continue;
            p.second.id = id++;
            blocks.push_back(p.first);
        }
        const int nb = (int)blocks.size();
        const int m = n_cell + 1;
        const bool sph_mesh = coord == KP_COORD_SPHERICAL;
        xf0.assign(3 * nb, 0.0);
        dx0.assign(3 * nb, 0.0);
        vol0.assign(nb, 0.0);
        xf.assign((size_t)3 * m * nb, 0.0);
        sph.assign(sph_mesh ? (size_t)KP_SPH_TAB(n_cell) * nb : 0, 0.0);
        level.assign(nb, 0);
        logic.assign(3 * nb, 0);
        nb_id.assign(24 * nb, -1);
        nb_shift.assign(24 * nb, 0);
        d_lvl.assign(6 * nb, 0);
        phybnd.assign(nb, 0);
        for (int b = 0; b < nb; ++b) {
            const reg_t &r = blocks[b];
            level[b] = r.level;
            for (int a = 0; a < 3; ++a) {
                logic[3 * b + a] = r.x[a];
                // set_coord_axes (meshgen.cpp:286-306): uniform axes are
                // xf0 + i*dx0, a non-uniform axis has dx0 = -1 and a table
                auto loc = [&](double i) {
                    const double xl = (r.x[a] + i / n_cell) / (double)(reg_max[a] << r.level);
                    return location(xl, a);
                };
                const bool uniform = !(a == 0 && dr_0 > 0.0);
                xf0[3 * b + a] = loc(0);
                dx0[3 * b + a] = uniform ? loc(1) - xf0[3 * b + a] : -1.0;
                double *f = &xf[((size_t)3 * b + a) * m];
                for (int k = 0; k <= n_cell; ++k)
                    f[k] = uniform ? xf0[3 * b + a] + k * dx0[3 * b + a] : loc(k);
            }
            if (sph_mesh) {
                // sph_t::volume factors (spherical.h:31-42) and the face
                // trigonometry of the spherical march; phi faces are reduced
                // mod 2 pi so the periodic seam is one plane in every block
                vol0[b] = -1.0;
                const int n = n_cell;
                double *t = &sph[(size_t)KP_SPH_TAB(n) * b];
                const double *fr = &xf[((size_t)3 * b + 0) * m];
                const double *ft = &xf[((size_t)3 * b + 1) * m];
                const double *fp = &xf[((size_t)3 * b + 2) * m];
                const double two_pi = 6.283185307179586476925286766559;
                for (int k = 0; k <= n; ++k) {
                    const double ph = std::fmod(fp[k], two_pi);
                    // the equatorial face is the plane z = 0 exactly
                    const double ct = std::cos(ft[k]);
                    t[k] = fr[k];
                    t[m + k] = std::fabs(ct) < 1e-15 ? 0.0 : ct;
                    t[2 * m + k] = std::sin(ph);
                    t[3 * m + k] = std::cos(ph);
                }
                for (int k = 0; k < n; ++k) {
                    t[4 * m + k] = (std::pow(fr[k + 1], 3.0) - std::pow(fr[k], 3.0)) / 3.0;
                    t[4 * m + n + k] = std::fabs(t[m + k] - t[m + k + 1]);
                    t[4 * m + 2 * n + k] = fp[k + 1] - fp[k];
                }
            } else {
                double v = 1.0;
                for (int a = 0; a < 3; ++a)
                    v *= dx0[3 * b + a];
                vol0[b] = v;
            }
            // associate_mesh phybnd (meshgen.cpp:390-398)
            const auto rl = region_max(r.level);
            for (int a = 0; a < 3; ++a)
                if (!periodic[a]) {
                    if (r.x[a] == 0)
                        phybnd[b] |= (uint8_t)(1u << (2 * a));
                    if (rl[a] - 1 == r.x[a])
                        phybnd[b] |= (uint8_t)(1u << (2 * a + 1));
                }
            // proxy_base_t::setup (proxy.h:22-62)
            for (const auto &f : neighbors(r)) {
                if (f.nb.mode != 0)
                    continue;
                for (int a = 0; a < 3; ++a) {
                    if (f.nb.x[a] >= 0 && f.nb.x[a] <= 1)
                        continue;
                    const int side = f.nb.x[a] > 0;
                    const int j0 = f.nb.x[(a + 1) % 3];
                    const int j1 = f.nb.x[(a + 2) % 3];
                    if (j0 < 0 || j0 > 1 || j1 < 0 || j1 > 1)
                        throw kp::error(KP_ERR_STATE, "face neighbour with bad sub-face offset");
                    const int slot = 24 * b + ((a * 2 + side) * 2 + j0) * 2 + j1;
                    nb_id[slot] = tree.at(f.guest).id;
                    d_lvl[6 * b + 2 * a + side] = (int8_t)f.nb.d_lvl;
                    const nbr_t rev = make_nbr(f.guest_raw, r);
                    uint8_t bits = 0;
                    for (int i = 0; i <= 1; ++i)
                        if (rev.x[(a + i + 1) % 3])
                            bits |= (uint8_t)(1u << i);
                    nb_shift[slot] = bits;
                }
            }
        }
        finalized = true;
    }
};

extern "C" {

kp_status kp_mesh_create(const int32_t root_blocks[3], int32_t n_cell, const double x_min[3],
                         const double x_max[3], const int32_t periodic[3], kp_mesh **out)
{
    return kp::guard([&] {
        if (!root_blocks || !x_min || !x_max || !out)
            throw kp::error(KP_ERR_ARG, "kp_mesh_create: null argument");
        if (n_cell < 2 || (n_cell & 1))
            throw kp::error(KP_ERR_ARG, "kp_mesh_create: n_cell must be even and >= 2");
        auto *m = new kp_mesh();
        for (int a = 0; a < 3; ++a) {
            if (root_blocks[a] < 1) {
                delete m;
                throw kp::error(KP_ERR_ARG, "kp_mesh_create: root_blocks must be >= 1");
            }
            if (!(x_max[a] > x_min[a])) {
                delete m;
                throw kp::error(KP_ERR_ARG, "kp_mesh_create: x_max must exceed x_min");
            }
            m->reg_max[a] = root_blocks[a];
            m->x_min[a] = x_min[a];
            m->x_max[a] = x_max[a];
            m->periodic[a] = periodic ? (periodic[a] != 0) : 0;
        }
        m->n_cell = n_cell;
        *out = m;
    });
}

void kp_mesh_destroy(kp_mesh *mesh) { delete mesh; }

kp_status kp_mesh_set_coord(kp_mesh *mesh, int32_t coord, double r_ratio)
{
    return kp::guard([&] {
        if (!mesh)
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: null mesh");
        if (mesh->finalized)
            throw kp::error(KP_ERR_STATE, "kp_mesh_set_coord: mesh already finalized");
        if (coord == KP_COORD_CARTESIAN) {
            mesh->coord = coord;
            mesh->ratio = 1.0;
            mesh->dr_0 = -1.0;
            return;
        }
        if (coord != KP_COORD_SPHERICAL)
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: unknown coordinate system");
        const double pi = 3.14159265358979323846, two_pi = 2.0 * pi;
        const auto &lo = mesh->x_min, &hi = mesh->x_max;
        if (!(lo[0] > 0.0))
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: spherical meshes need r_min > 0");
        if (!(lo[1] >= 0.0 && hi[1] <= pi + 1e-12))
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: theta limits must lie in [0, pi]");
        const double wphi = hi[2] - lo[2];
        if (!(wphi <= two_pi + 1e-12))
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: phi range wider than 2 pi");
        if (!(wphi / (mesh->reg_max[2] * mesh->n_cell) < pi))
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: phi cells must be narrower than pi");
        if (mesh->periodic[0] || mesh->periodic[1])
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: only phi may be periodic");
        if (mesh->periodic[2] && std::fabs(wphi - two_pi) > 1e-9)
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: a periodic phi axis spans 2 pi");
        if (!std::isfinite(r_ratio) || r_ratio == 0.0)
            throw kp::error(KP_ERR_ARG, "kp_mesh_set_coord: bad r_ratio");
        // cyl_t::read (cylindrical.h:18-32)
        const int n_r = mesh->reg_max[0] * mesh->n_cell;
        double q = r_ratio, dr = -1.0;
        if (std::fabs(q - 1.0) >= 1e-6) {
            if (q < 0.0)
                q = std::pow(hi[0] / lo[0], 1.0 / n_r);
            dr = (hi[0] - lo[0]) / ((std::pow(q, n_r) - 1.0) / (q - 1.0));
        } else {
            q = 1.0;
        }
        mesh->coord = coord;
        mesh->ratio = q;
        mesh->dr_0 = dr;
    });
}

kp_status kp_mesh_refine(kp_mesh *mesh, int32_t level, const int32_t *idx, int64_t n)
{
    return kp::guard([&] {
        if (!mesh || (n > 0 && !idx))
            throw kp::error(KP_ERR_ARG, "kp_mesh_refine: null argument");
        if (mesh->finalized)
            throw kp::error(KP_ERR_STATE, "kp_mesh_refine: mesh already finalized");
        if (level < 0 || level > 20)
            throw kp::error(KP_ERR_ARG, "kp_mesh_refine: level out of range");
        for (int64_t k = 0; k < n; ++k) {
            reg_t r;
            r.level = level;
            r.x = {idx[3 * k], idx[3 * k + 1], idx[3 * k + 2]};
            if (!mesh->is_inside(r))
                throw kp::error(KP_ERR_ARG, "kp_mesh_refine: region outside the domain");
            mesh->insert(r);
        }
    });
}

kp_status kp_mesh_refine_region(kp_mesh *mesh, int32_t level, const double x_lo[3],
                                const double x_hi[3])
{
    return kp::guard([&] {
        if (!mesh || !x_lo || !x_hi)
            throw kp::error(KP_ERR_ARG, "kp_mesh_refine_region: null argument");
        if (mesh->finalized)
            throw kp::error(KP_ERR_STATE, "kp_mesh_refine_region: mesh already finalized");
        if (level <= 0)
            return;
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = mesh->search_blk_idx(x_lo[a], level, a);
            hi[a] = mesh->search_blk_idx(x_hi[a], level, a);
            if (lo[a] < 0 || hi[a] < 0)
                throw kp::error(KP_ERR_ARG, "Incorrect SMR region");
        }
        for (int i2 = lo[2]; i2 <= hi[2]; ++i2)
            for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
                for (int i0 = lo[0]; i0 <= hi[0]; ++i0) {
                    reg_t r;
                    r.level = level;
                    r.x = {i0, i1, i2};
                    mesh->insert(r);
                }
    });
}

kp_status kp_mesh_finalize(kp_mesh *mesh)
{
    return kp::guard([&] {
        if (!mesh)
            throw kp::error(KP_ERR_ARG, "kp_mesh_finalize: null mesh");
        if (mesh->finalized)
            throw kp::error(KP_ERR_STATE, "kp_mesh_finalize: already finalized");
        mesh->finalize();
    });
}

kp_status kp_mesh_describe(const kp_mesh *mesh, kp_mesh_desc *d)
{
    return kp::guard([&] {
        if (!mesh || !d)
            throw kp::error(KP_ERR_ARG, "kp_mesh_describe: null argument");
        if (!mesh->finalized)
            throw kp::error(KP_ERR_STATE, "kp_mesh_describe: mesh not finalized");
        std::memset(d, 0, sizeof(*d));
        d->n_blocks = (int32_t)mesh->blocks.size();
        d->n_cell = mesh->n_cell;
        for (int a = 0; a < 3; ++a) {
            d->periodic[a] = mesh->periodic[a];
            d->x_min[a] = mesh->x_min[a];
            d->x_max[a] = mesh->x_max[a];
        }
        d->xf0 = mesh->xf0.data();
        d->dx0 = mesh->dx0.data();
        d->vol0 = mesh->vol0.data();
        d->level = mesh->level.data();
        d->logic = mesh->logic.data();
        d->nb_id = mesh->nb_id.data();
        d->nb_shift = mesh->nb_shift.data();
        d->d_lvl = mesh->d_lvl.data();
        d->phybnd = mesh->phybnd.data();
        d->coord = mesh->coord;
        d->xf = mesh->xf.data();
        d->sph = mesh->coord == KP_COORD_SPHERICAL ? mesh->sph.data() : nullptr;
    });
}

kp_status kp_mesh_locate(const kp_mesh *mesh, const double x[3], int32_t *ib, int32_t i[3])
{
    return kp::guard([&] {
        if (!mesh || !x || !ib || !i)
            throw kp::error(KP_ERR_ARG, "kp_mesh_locate: null argument");
        if (!mesh->finalized)
            throw kp::error(KP_ERR_STATE, "kp_mesh_locate: mesh not finalized");
        *ib = -1;
        i[0] = i[1] = i[2] = 0;
        const int n = mesh->n_cell;
        if (mesh->coord != KP_COORD_CARTESIAN) {
            // geo_t::idx over the face tables: the cell k with xf[k] <= x < xf[k+1]
            const int m = n + 1;
            for (size_t b = 0; b < mesh->blocks.size(); ++b) {
                int c[3];
                bool in = true;
                for (int a = 0; a < 3 && in; ++a) {
                    const double *f = &mesh->xf[(3 * b + a) * m];
                    in = (x[a] >= f[0] && x[a] < f[n]);
                    int k = 0;
                    while (in && k + 1 < n && f[k + 1] <= x[a])
                        ++k;
                    c[a] = k;
                }
                if (!in)
                    continue;
                *ib = (int32_t)b;
                for (int a = 0; a < 3; ++a)
                    i[a] = c[a];
                return;
            }
            return;
        }
        for (size_t b = 0; b < mesh->blocks.size(); ++b) {
            bool in = true;
            for (int a = 0; a < 3 && in; ++a) {
                const double lo = mesh->xf0[3 * b + a];
                const double hi = lo + n * mesh->dx0[3 * b + a];
                in = (x[a] >= lo && x[a] < hi);
            }
            if (!in)
                continue;
            *ib = (int32_t)b;
            for (int a = 0; a < 3; ++a) {
                int c = (int)((x[a] - mesh->xf0[3 * b + a]) / mesh->dx0[3 * b + a]);
                i[a] = std::min(std::max(c, 0), n - 1);
            }
            return;
        }
    });
}

} // extern "C"
