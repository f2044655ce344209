// kp_module.cpp -- the photon module: host-side counterpart of the
// reference's particle::radiation::base_t (radiation.h:19-55 over
// particle::base_t, driver.h:15-102, driver.cpp:1-73), written against the
// engine's C ABI (kpolrad.h) exactly as a reference-side caller would be.
//
//   read  : INI text with the reference's section|key map (io/args/input.cpp
//           read_ascii: [section] headers, key = value, '#' comments)
//   init  : meshgen root grid + refine_region boxes + 2:1 tree (kp_mesh_*),
//           engine on the device, uniform [photon] medium
//   step  : n_cycle_stp cycles of kp_forward, fresh packet ids per cycle
//   save  : binary_io container (io/binary/binary_io.h:6-32 format,
//           utilities/data_transfer/serialize.h encodings) with the rad_t
//           tags of rad_data.cpp:45-49 / dat_3d.h:117-126 per block.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/kp_module.h"
#include "kp_error.h"

namespace {

#define KP_HIPM(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw kp::error(KP_ERR_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

std::string trim(const std::string &s)
{
    const auto b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos)
        return "";
    const auto e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

// The reference's `input`: a map "section|key" -> value text.
struct ini_t {
    std::map<std::string, std::string> items;
    std::vector<std::string> sections;

    void parse(const std::string &text)
    {
        std::istringstream in(text);
        std::string line, sec;
        while (std::getline(in, line)) {
            line = trim(line.substr(0, line.find('#')));
            if (line.empty())
                continue;
            if (line[0] == '[') {
                const auto e = line.find(']');
                if (e == std::string::npos)
                    throw kp::error(KP_ERR_ARG, "ini: incorrect input section: " + line);
                sec = line.substr(1, e - 1);
                if (std::find(sections.begin(), sections.end(), sec) == sections.end())
                    sections.push_back(sec);
                continue;
            }
            const auto eq = line.find('=');
            const std::string key = trim(line.substr(0, eq));
            const std::string val = eq == std::string::npos ? "" : trim(line.substr(eq + 1));
            items[(sec.empty() ? "" : sec + "|") + key] = val;
        }
    }
    bool has(const std::string &s, const std::string &k) const
    {
        return items.count(s + "|" + k) != 0;
    }
    template <class T> std::vector<T> list(const std::string &s, const std::string &k) const
    {
        std::vector<T> v;
        auto it = items.find(s + "|" + k);
        if (it == items.end())
            return v;
        std::istringstream ss(it->second);
        T x;
        while (ss >> x)
            v.push_back(x);
        return v;
    }
    template <class T> T get(const std::string &s, const std::string &k, T def) const
    {
        auto v = list<T>(s, k);
        return v.empty() ? def : v[0];
    }
    template <class T> T need(const std::string &s, const std::string &k) const
    {
        auto v = list<T>(s, k);
        if (v.empty())
            throw kp::error(KP_ERR_ARG, "ini: missing " + s + "|" + k);
        return v[0];
    }
    template <class T, size_t N>
    std::array<T, N> arr(const std::string &s, const std::string &k, bool required,
                         std::array<T, N> def) const
    {
        auto v = list<T>(s, k);
        if (v.empty()) {
            if (required)
                throw kp::error(KP_ERR_ARG, "ini: missing " + s + "|" + k);
            return def;
        }
        if (v.size() < N)
            throw kp::error(KP_ERR_ARG, "ini: " + s + "|" + k + " needs " + std::to_string(N) +
                                            " values");
        std::array<T, N> a;
        std::copy(v.begin(), v.begin() + N, a.begin());
        return a;
    }
};

// ------------------------------------------------------------ binary_io
// Container of io/binary/binary_io.h:6-32: one byte (is_little | sizeof
// size_t), size_t header size, the serialized map tag -> {size, u_size,
// offset} (size_t count; per entry size_t length + tag bytes, size_t size,
// char u_size, size_t offset; map order), then the bulk data at
// header + offset, offsets cumulative in map order.
struct blob_t {
    std::vector<unsigned char> data;
    char u_size = 1;
};

template <class T> void put(std::string &s, const T &v)
{
    s.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

void write_container(const std::string &path, const std::map<std::string, blob_t> &tags)
{
    std::string hdr;
    put<size_t>(hdr, tags.size());
    size_t off = 0;
    for (const auto &p : tags) {
        put<size_t>(hdr, p.first.size());
        hdr.append(p.first);
        put<size_t>(hdr, p.second.data.size());
        put<char>(hdr, p.second.u_size);
        put<size_t>(hdr, off);
        off += p.second.data.size();
    }
    const int one = 1;
    const char is_le = (*reinterpret_cast<const char *>(&one) == 1) ? 1 : 0;
    const char lead = (char)(is_le | (char)sizeof(size_t));
    const size_t header_size = 1 + sizeof(size_t) + hdr.size();
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f)
        throw kp::error(KP_ERR_ARG, "kp_module_save: cannot open " + path);
    f.write(&lead, 1);
    f.write(reinterpret_cast<const char *>(&header_size), sizeof(size_t));
    f.write(hdr.data(), (std::streamsize)hdr.size());
    for (const auto &p : tags)
        f.write(reinterpret_cast<const char *>(p.second.data.data()),
                (std::streamsize)p.second.data.size());
    if (!f)
        throw kp::error(KP_ERR_ARG, "kp_module_save: write failed: " + path);
}

template <class T> blob_t make_blob(const T *p, size_t count)
{
    blob_t b;
    b.u_size = (char)sizeof(T);
    b.data.resize(sizeof(T) * count);
    if (count)
        std::memcpy(b.data.data(), p, sizeof(T) * count);
    return b;
}

// serialize::write of std::map<std::string, std::string> (input_args)
blob_t map_blob(const std::map<std::string, std::string> &m)
{
    std::string s;
    put<size_t>(s, m.size());
    for (const auto &p : m) {
        put<size_t>(s, p.first.size());
        s.append(p.first);
        put<size_t>(s, p.second.size());
        s.append(p.second);
    }
    blob_t b;
    b.u_size = 1;
    b.data.assign(s.begin(), s.end());
    return b;
}

template <class T> struct dbuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        free();
        KP_HIPM(hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1)));
        KP_HIPM(hipMemset(p, 0, sizeof(T) * std::max<size_t>(count, 1)));
        n = count;
    }
    void free()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~dbuf() { free(); }
};

} // namespace

struct kp_module {
    ini_t ini;
    // [mesh]
    std::array<double, 3> x_min{0, 0, 0}, x_max{1, 1, 1};
    std::array<int, 3> n_global{0, 0, 0}, n_block{0, 0, 0};
    std::array<int, 3> periodic{0, 0, 0};
    int coord = KP_COORD_CARTESIAN;
    double r_ratio = 1.0;
    bool sched_given = false;
    int order = -1; // photon|order: -1 = the engine's default, 0 slot, 1 direction
    // [particle]
    int n_hv = 1, n_cyc = 1, n_th = 64, step_max = 100;
    int64_t n_par = 1;
    // [photon]
    uint32_t seed = 42;
    int model = KP_THOMSON, precision = KP_F64, sched_kind = KP_SCHED_PERSISTENT;
    double p_max = 1.0, kappa = 0.0, weight = 1.0, rr_weight = 1e-2, rr_survive = 0.1;
    std::vector<double> albedo, g, kappa_scale, incl;
    std::array<double, 3> source{0, 0, 0};
    int image_n = 128;
    double image_half = 0.5;

    kp_mesh *mesh = nullptr;
    kp_engine *eng = nullptr;
    int device = -1;
    int64_t n_cells = 0;
    int n_cell = 0;
    int64_t packets_done = 0;
    std::vector<float> kappa_h; // medium copy kept for save()
    dbuf<double> J, image;
    dbuf<uint64_t> counters;

    ~kp_module()
    {
        J.free();
        image.free();
        counters.free();
        if (eng)
            kp_engine_destroy(eng);
        if (mesh)
            kp_mesh_destroy(mesh);
    }

    // particle::radiation::base_t::read and the drivers' init reads
    void read(const std::string &text)
    {
        ini.parse(text);
        x_min = ini.arr<double, 3>("mesh", "x_min", true, {});
        x_max = ini.arr<double, 3>("mesh", "x_max", true, {});
        n_global = ini.arr<int, 3>("mesh", "n_cell_global", true, {});
        n_block = ini.arr<int, 3>("mesh", "n_cell_block", false, n_global);
        for (int a = 0; a < 3; ++a) {
            if (n_global[a] <= 0 || n_block[a] <= 0 || n_global[a] % n_block[a])
                throw kp::error(KP_ERR_ARG, "ini: mesh size indivisible by sub-mesh");
            if (!(x_max[a] > x_min[a]))
                throw kp::error(KP_ERR_ARG, "ini: mesh|x_max must exceed x_min");
        }
        if (n_block[0] != n_block[1] || n_block[0] != n_block[2] || (n_block[0] & 1))
            throw kp::error(KP_ERR_UNSUPPORTED, "ini: cubic blocks of an even cell count only");
        // meshgen::sph_t (spherical.h) and its [mesh] r_ratio (cylindrical.h:21-23)
        const std::string cs = ini.get<std::string>("mesh", "coord", "cartesian");
        if (cs == "spherical")
            coord = KP_COORD_SPHERICAL;
        else if (cs != "cartesian")
            throw kp::error(KP_ERR_ARG, "ini: mesh|coord must be cartesian or spherical");
        r_ratio = ini.get<double>("mesh", "r_ratio", 1.0);
        const auto kinds = ini.list<std::string>("boundary", "kinds");
        for (size_t s = 0; s < kinds.size() && s < 6; ++s)
            if (kinds[s].find("per") != std::string::npos)
                periodic[s / 2] = 1;
        n_hv = ini.get<int>("particle", "n_hv", 1);
        n_cyc = ini.get<int>("particle", "n_cycle_stp", 1);
        n_th = ini.get<int>("particle", "n_th", 64);
        step_max = ini.get<int>("particle", "step_max", 100);
        n_par = (int64_t)std::llround(ini.get<double>("particle", "n_par", 1.0));
        if (n_hv < 1 || n_cyc < 0 || n_par < 0)
            throw kp::error(KP_ERR_ARG, "ini: particle|n_hv, n_cycle_stp, n_par out of range");
        seed = (uint32_t)ini.get<long long>("photon", "seed", 42);
        const std::string m = ini.get<std::string>("photon", "model", "thomson");
        if (m == "thomson")
            model = KP_THOMSON;
        else if (m == "hg_rayleigh" || m == "hg")
            model = KP_HG_RAYLEIGH;
        else
            throw kp::error(KP_ERR_ARG, "ini: photon|model must be thomson or hg_rayleigh");
        const std::string pr = ini.get<std::string>("photon", "precision", "f64");
        precision = pr == "f32" ? KP_F32 : KP_F64;
        const std::string sk = ini.get<std::string>("photon", "schedule", "persistent");
        if (sk == "history")
            sched_kind = KP_SCHED_HISTORY;
        else if (sk == "persistent")
            sched_kind = KP_SCHED_PERSISTENT;
        else if (sk == "block_queue")
            sched_kind = KP_SCHED_BLOCKQ;
        else
            throw kp::error(KP_ERR_ARG,
                            "ini: photon|schedule must be persistent, history or block_queue");
        sched_given = ini.has("photon", "schedule");
        if (ini.has("photon", "order")) {
            const std::string o = ini.get<std::string>("photon", "order", "slot");
            if (o == "slot")
                order = 0;
            else if (o == "direction")
                order = 1;
            else
                throw kp::error(KP_ERR_ARG, "ini: photon|order must be slot or direction");
        }
        p_max = ini.get<double>("photon", "p_max", 1.0);
        kappa = ini.get<double>("photon", "kappa", 0.0);
        weight = ini.get<double>("photon", "weight", 1.0);
        rr_weight = ini.get<double>("photon", "rr_weight", 1e-2);
        rr_survive = ini.get<double>("photon", "rr_survive", 0.1);
        auto per_band = [&](const char *k, double def) {
            auto v = ini.list<double>("photon", k);
            if (v.empty())
                v.push_back(def);
            if ((int)v.size() == 1)
                v.resize(n_hv, v[0]);
            if ((int)v.size() != n_hv)
                throw kp::error(KP_ERR_ARG, std::string("ini: photon|") + k +
                                                " needs 1 or n_hv values");
            return v;
        };
        albedo = per_band("albedo", 1.0);
        g = per_band("g", 0.0);
        // the spherical photon streams (photon_sph.h): no extinction by default
        kappa_scale = per_band("kappa_scale", coord == KP_COORD_SPHERICAL ? 0.0 : 1.0);
        incl = ini.list<double>("photon", "incl");
        for (int a = 0; a < 3; ++a)
            source[a] = 0.5 * (x_min[a] + x_max[a]);
        if (coord == KP_COORD_SPHERICAL) {
            // the centre of the (r, theta, phi) box as a Cartesian point
            const double r = source[0], th = source[1], ph = source[2];
            source = {r * std::sin(th) * std::cos(ph), r * std::sin(th) * std::sin(ph),
                      r * std::cos(th)};
            if (!incl.empty() || precision != KP_F64)
                throw kp::error(KP_ERR_UNSUPPORTED, "ini: spherical meshes stream in f64 "
                                                    "without observers (photon|incl)");
        }
        source = ini.arr<double, 3>("photon", "source", false, source);
        image_n = ini.get<int>("photon", "image_n", 128);
        image_half = ini.get<double>("photon", "image_half", 0.5);
    }

    void init(int dev)
    {
        if (mesh)
            throw kp::error(KP_ERR_STATE, "kp_module_init: already initialised");
        int32_t root[3];
        for (int a = 0; a < 3; ++a)
            root[a] = n_global[a] / n_block[a];
        n_cell = n_block[0];
        int32_t per[3] = {periodic[0], periodic[1], periodic[2]};
        kp_status st = kp_mesh_create(root, n_cell, x_min.data(), x_max.data(), per, &mesh);
        if (st != KP_OK)
            throw kp::error(st, kp_last_error());
        if (coord != KP_COORD_CARTESIAN) {
            st = kp_mesh_set_coord(mesh, coord, r_ratio);
            if (st != KP_OK)
                throw kp::error(st, kp_last_error());
        }
        // load_refinement: every [refine_region*] section, in file order
        for (const auto &sec : ini.sections) {
            if (sec.find("refine_region") == std::string::npos)
                continue;
            const int level = ini.get<int>(sec, "level", 0);
            std::array<double, 3> lo = ini.arr<double, 3>(sec, "loc", false, {0, 0, 0});
            lo = ini.arr<double, 3>(sec, "x_min", false, lo);
            std::array<double, 3> hi = ini.arr<double, 3>(sec, "x_max", false, lo);
            st = kp_mesh_refine_region(mesh, level, lo.data(), hi.data());
            if (st != KP_OK)
                throw kp::error(st, kp_last_error());
        }
        st = kp_mesh_finalize(mesh);
        if (st != KP_OK)
            throw kp::error(st, kp_last_error());
        kp_mesh_desc d;
        kp_mesh_describe(mesh, &d);
        n_cells = (int64_t)d.n_blocks * n_cell * n_cell * n_cell;
        kappa_h.assign((size_t)n_cells, (float)kappa);
        if (dev < 0)
            return;
        device = dev;
        KP_HIPM(hipSetDevice(dev));
        st = kp_engine_create(&d, dev, &eng);
        if (st != KP_OK)
            throw kp::error(st, kp_last_error());
        kp_schedule s;
        kp_get_schedule(eng, &s);
        // spherical engines keep their own default unless the input names one
        if (coord == KP_COORD_CARTESIAN || sched_given)
            s.kind = sched_kind;
        if (s.kind == KP_SCHED_HISTORY)
            s.block_threads = 256;
        if (order >= 0)
            s.coherent = order;
        st = kp_set_schedule(eng, &s);
        if (st != KP_OK)
            throw kp::error(st, kp_last_error());
        std::vector<float> a((size_t)n_cells, (float)albedo[0]), gg((size_t)n_cells, (float)g[0]);
        set_media(kappa_h.data(), a.data(), gg.data());
        J.alloc((size_t)n_cells * n_hv);
        if (!incl.empty())
            image.alloc((size_t)n_hv * incl.size() * image_n * image_n * 3);
        counters.alloc(KP_NCOUNTERS);
    }

    void set_media(const float *k, const float *a, const float *gg)
    {
        if (!eng)
            throw kp::error(KP_ERR_STATE, "kp_module_set_media: kp_module_init with a device first");
        kp_status st = kp_set_media(eng, k, a, gg);
        if (st != KP_OK)
            throw kp::error(st, kp_last_error());
        kappa_h.assign(k, k + n_cells);
    }

    void step(hipStream_t stream)
    {
        if (!eng)
            throw kp::error(KP_ERR_STATE, "kp_module_step: kp_module_init with a device first");
        std::vector<kp_band> bands((size_t)n_hv);
        for (int b = 0; b < n_hv; ++b) {
            bands[b].kappa_scale = kappa_scale[b];
            bands[b].albedo = albedo[b];
            bands[b].g = g[b];
            bands[b].n_packets = (uint64_t)n_par;
            bands[b].flags = 0;
            bands[b].pad_ = 0;
        }
        kp_source src;
        for (int a = 0; a < 3; ++a)
            src.x[a] = source[a];
        src.weight = weight;
        kp_run r{};
        r.seed = seed;
        r.scatter_model = model;
        r.p_max = p_max;
        r.n_bands = n_hv;
        r.precision = precision;
        r.bands = bands.data();
        r.step_max = step_max;
        r.n_incl = (int32_t)incl.size();
        r.incl_deg = incl.empty() ? nullptr : incl.data();
        r.rr_weight = rr_weight;
        r.rr_survive = rr_survive;
        r.image_n = image_n;
        r.image_half = image_half;
        kp_tallies t{};
        t.J = J.p;
        t.image = image.p;
        t.counters = counters.p;
        for (int c = 0; c < n_cyc; ++c) {
            r.packet_offset = (uint64_t)packets_done;
            const kp_status st = kp_forward(eng, &src, &r, &t, stream);
            if (st != KP_OK)
                throw kp::error(st, kp_last_error());
            packets_done += n_par;
        }
        KP_HIPM(hipStreamSynchronize(stream));
    }

    void tallies(double *Jh, double *img, uint64_t *cnt) const
    {
        if (Jh) {
            if (J.p)
                KP_HIPM(hipMemcpy(Jh, J.p, sizeof(double) * (size_t)n_cells * n_hv,
                                  hipMemcpyDeviceToHost));
            else
                std::fill(Jh, Jh + (size_t)n_cells * n_hv, 0.0);
        }
        const size_t ni = (size_t)n_hv * incl.size() * image_n * image_n * 3;
        if (img) {
            if (image.p)
                KP_HIPM(hipMemcpy(img, image.p, sizeof(double) * ni, hipMemcpyDeviceToHost));
            else
                std::fill(img, img + ni, 0.0);
        }
        if (cnt) {
            if (counters.p)
                KP_HIPM(hipMemcpy(cnt, counters.p, sizeof(uint64_t) * KP_NCOUNTERS,
                                  hipMemcpyDeviceToHost));
            else
                std::fill(cnt, cnt + KP_NCOUNTERS, 0ull);
        }
    }

    void save(const std::string &path) const
    {
        if (!mesh)
            throw kp::error(KP_ERR_STATE, "kp_module_save: kp_module_init first");
        kp_mesh_desc d;
        kp_mesh_describe(mesh, &d);
        const size_t n3 = (size_t)n_cell * n_cell * n_cell;
        std::vector<double> Jh((size_t)n_cells * n_hv);
        const size_t ni = (size_t)n_hv * incl.size() * image_n * image_n * 3;
        std::vector<double> img(ni);
        std::vector<uint64_t> cnt(KP_NCOUNTERS);
        tallies(Jh.data(), ni ? img.data() : nullptr, cnt.data());
        std::map<std::string, blob_t> tags;
        tags["input_args"] = map_blob(ini.items);
        const int n_int = n_hv, n_step = 1, one = 1;
        const int ncell3[3] = {n_cell, n_cell, n_cell}, zeros3[3] = {0, 0, 0};
        for (int b = 0; b < d.n_blocks; ++b) {
            const std::string pre = "block_" + std::to_string(b) + "|";
            tags[pre + "i_logic"] = make_blob(d.logic + 3 * b, 3);
            tags[pre + "level"] = make_blob(d.level + b, 1);
            const std::string f = pre + "rad_flx_flux_";
            tags[f + "n_int"] = make_blob(&n_int, 1);
            tags[f + "n_step"] = make_blob(&n_step, 1);
            tags[f + "n_cell"] = make_blob(ncell3, 3);
            tags[f + "n_gh"] = make_blob(zeros3, 3);
            tags[f + "field"] = make_blob(Jh.data() + (size_t)b * n3 * n_hv, n3 * n_hv);
            const std::string k = pre + "rad_flx_mfpi_";
            tags[k + "n_int"] = make_blob(&one, 1);
            tags[k + "n_step"] = make_blob(&n_step, 1);
            tags[k + "n_cell"] = make_blob(ncell3, 3);
            tags[k + "n_gh"] = make_blob(zeros3, 3);
            std::vector<double> kd(kappa_h.begin() + (ptrdiff_t)(b * n3),
                                   kappa_h.begin() + (ptrdiff_t)((b + 1) * n3));
            tags[k + "field"] = make_blob(kd.data(), n3);
        }
        if (ni) {
            tags["photon|image"] = make_blob(img.data(), ni);
            const int shape[5] = {n_hv, (int)incl.size(), image_n, image_n, 3};
            tags["photon|image_shape"] = make_blob(shape, 5);
        }
        tags["photon|counters"] = make_blob(cnt.data(), cnt.size());
        tags["photon|packets_per_band"] = make_blob(&packets_done, 1);
        write_container(path, tags);
    }
};

extern "C" {

kp_status kp_module_create(const char *ini_text, kp_module **out)
{
    return kp::guard([&] {
        if (!ini_text || !out)
            throw kp::error(KP_ERR_ARG, "kp_module_create: null argument");
        auto m = std::make_unique<kp_module>();
        m->read(ini_text);
        *out = m.release();
    });
}

void kp_module_destroy(kp_module *m) { delete m; }

kp_status kp_module_init(kp_module *m, int32_t device)
{
    return kp::guard([&] {
        if (!m)
            throw kp::error(KP_ERR_ARG, "kp_module_init: null module");
        m->init(device);
    });
}

kp_status kp_module_set_media(kp_module *m, const float *kappa_rho, const float *albedo,
                              const float *g)
{
    return kp::guard([&] {
        if (!m || !kappa_rho || !albedo)
            throw kp::error(KP_ERR_ARG, "kp_module_set_media: null argument");
        m->set_media(kappa_rho, albedo, g);
    });
}

kp_status kp_module_step(kp_module *m, void *stream)
{
    return kp::guard([&] {
        if (!m)
            throw kp::error(KP_ERR_ARG, "kp_module_step: null module");
        m->step((hipStream_t)stream);
    });
}

kp_status kp_module_save(const kp_module *m, const char *path)
{
    return kp::guard([&] {
        if (!m || !path)
            throw kp::error(KP_ERR_ARG, "kp_module_save: null argument");
        m->save(path);
    });
}

kp_status kp_module_get_info(const kp_module *m, kp_module_info *info)
{
    return kp::guard([&] {
        if (!m || !info)
            throw kp::error(KP_ERR_ARG, "kp_module_get_info: null argument");
        kp_module_info i{};
        i.n_cell = m->n_cell;
        i.n_cells = m->n_cells;
        i.n_blocks = m->n_cell ? m->n_cells / ((int64_t)m->n_cell * m->n_cell * m->n_cell) : 0;
        i.n_bands = m->n_hv;
        i.n_incl = (int32_t)m->incl.size();
        i.image_n = m->image_n;
        i.n_par = m->n_par;
        i.n_cycle = m->n_cyc;
        i.initialised = m->mesh != nullptr;
        i.packets_done = m->packets_done;
        *info = i;
    });
}

kp_status kp_module_tallies(const kp_module *m, double *J, double *image, uint64_t *counters)
{
    return kp::guard([&] {
        if (!m)
            throw kp::error(KP_ERR_ARG, "kp_module_tallies: null module");
        if (!m->mesh)
            throw kp::error(KP_ERR_STATE, "kp_module_tallies: kp_module_init first");
        m->tallies(J, image, counters);
    });
}

kp_status kp_module_mesh(const kp_module *m, const kp_mesh **mesh)
{
    return kp::guard([&] {
        if (!m || !mesh)
            throw kp::error(KP_ERR_ARG, "kp_module_mesh: null argument");
        if (!m->mesh)
            throw kp::error(KP_ERR_STATE, "kp_module_mesh: kp_module_init first");
        *mesh = m->mesh;
    });
}

} // extern "C"
