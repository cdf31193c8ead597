// grt_io.cpp -- the file formats either side of the traced path, host only:
//
//   grt_load_config        the flat TOML subset, key set, defaults and checks of
//                          toml_config.hpp:40-277 (defaults utils.cpp:4-37,
//                          utils.hpp:185-186, simulation_options.hpp:13-32,
//                          photon_generation.hpp:40-75)
//   grt_read_npy           f8 C-order .npy reader (cnpy.h:302-349 header rules,
//                          load_python_arrays.hpp:82-112 messages)
//   grt_load_numpy_fields  the seven-file bundle of load_python_arrays.hpp:15-223
//                          handed to grt_set_grid
//   grt_write_*_npz        uncompressed ZIP archives of .npy members, published
//                          atomically through "<path>.tmp" + rename
//                          (write_output.hpp:66-142, cnpy.h:94-300)
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "grt.h"

// The library's message slot for calls without a context (grt_api.cpp).
void grt_set_global_error(const std::string& msg);
int grt_set_error(grt_ctx* ctx, int code, const std::string& msg);

namespace {

const double kPi = 3.14159265358979323846;

std::string strip(const std::string& s)
{
    const auto b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    const auto e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

using RawConfig = std::map<std::string, std::string>;

// toml_config.hpp:40-77
RawConfig parse_flat_toml(const std::string& path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open config file: " + path);
    RawConfig values;
    std::string section, line;
    while (std::getline(in, line)) {
        bool quoted = false;
        for (size_t i = 0; i < line.size(); ++i) {
            if (line[i] == '"') quoted = !quoted;
            else if (line[i] == '#' && !quoted) { line = line.substr(0, i); break; }
        }
        const std::string t = strip(line);
        if (t.empty()) continue;
        if (t.front() == '[' && t.back() == ']') {
            section = strip(t.substr(1, t.size() - 2));
            continue;
        }
        const auto eq = t.find('=');
        if (eq == std::string::npos)
            throw std::runtime_error("Malformed config line (expected 'key = value'): " + line);
        const std::string key = strip(t.substr(0, eq));
        std::string value = strip(t.substr(eq + 1));
        if (value.size() >= 2 && value.front() == '"' && value.back() == '"')
            value = value.substr(1, value.size() - 2);
        values[section.empty() ? key : section + "." + key] = value;
    }
    return values;
}

// toml_config.hpp:79-123 (std::stod / stoi / stoull semantics)
double get_real(const RawConfig& c, const std::string& k, double fb)
{
    const auto it = c.find(k);
    return it == c.end() ? fb : std::stod(it->second);
}
int get_int(const RawConfig& c, const std::string& k, int fb)
{
    const auto it = c.find(k);
    return it == c.end() ? fb : std::stoi(it->second);
}
uint64_t get_unsigned(const RawConfig& c, const std::string& k, uint64_t fb)
{
    const auto it = c.find(k);
    if (it == c.end()) return fb;
    if (!it->second.empty() && it->second.front() == '-')
        throw std::runtime_error(k + " must not be negative");
    return std::stoull(it->second);
}
bool get_bool(const RawConfig& c, const std::string& k, bool fb)
{
    const auto it = c.find(k);
    if (it == c.end()) return fb;
    if (it->second == "true") return true;
    if (it->second == "false") return false;
    throw std::runtime_error(k + " must be true or false");
}
std::string get_string(const RawConfig& c, const std::string& k, const std::string& fb)
{
    const auto it = c.find(k);
    return it == c.end() ? fb : it->second;
}

void copy_string(char* dst, size_t cap, const std::string& s, const char* key)
{
    if (s.size() + 1 > cap) throw std::runtime_error(std::string(key) + " is too long");
    std::memcpy(dst, s.c_str(), s.size() + 1);
}

// output_variables.hpp:11-96: the normalised comma list of a selection
const char* const kVariables[] = {"id", "frequency", "emission_frame_energy", "x0", "x1", "x2",
                                  "x3", "k0", "k1", "k2", "k3", "I", "Q", "U", "V", "dlambda",
                                  "terminate", "phase", "theta_disp", "phi_disp"};

std::vector<std::string> parse_selection(std::string spec)
{
    spec = strip(spec);
    if (spec.empty()) throw std::runtime_error("output variables cannot be empty");
    if (spec.front() == '[' && spec.back() == ']') spec = spec.substr(1, spec.size() - 2);
    std::vector<std::string> names;
    size_t begin = 0;
    while (begin <= spec.size()) {
        const size_t comma = spec.find(',', begin);
        std::string name = strip(spec.substr(begin, comma == std::string::npos ? std::string::npos
                                                                                : comma - begin));
        if (name.size() >= 2 && ((name.front() == '"' && name.back() == '"') ||
                                 (name.front() == '\'' && name.back() == '\'')))
            name = strip(name.substr(1, name.size() - 2));
        if (name == "all") {
            names.assign(std::begin(kVariables), std::end(kVariables));
            return names;
        }
        if (!name.empty()) {
            if (std::find(std::begin(kVariables), std::end(kVariables), name) == std::end(kVariables))
                throw std::runtime_error("unknown output variable: " + name);
            if (std::find(names.begin(), names.end(), name) == names.end()) names.push_back(name);
        }
        if (comma == std::string::npos) break;
        begin = comma + 1;
    }
    if (names.empty()) throw std::runtime_error("output variables cannot be empty");
    return names;
}

std::string join(const std::vector<std::string>& v)
{
    std::string s;
    for (const auto& x : v) s += (s.empty() ? "" : ",") + x;
    return s;
}

// ---- .npy (NumPy format 1.0 - 3.0) ----------------------------------------

struct NpyHeader {
    std::string descr;
    bool fortran = false;
    std::vector<uint64_t> shape;
    uint64_t data_offset = 0;
};

NpyHeader read_npy_header(std::ifstream& in, const std::string& name)
{
    char magic[8];
    in.read(magic, 8);
    if (!in || std::memcmp(magic, "\x93NUMPY", 6) != 0)
        throw std::runtime_error("Not a NumPy .npy file: " + name);
    const int major = (unsigned char)magic[6];
    uint32_t hlen = 0;
    if (major == 1) {
        unsigned char b[2];
        in.read(reinterpret_cast<char*>(b), 2);
        hlen = b[0] | (b[1] << 8);
    } else if (major == 2 || major == 3) {
        unsigned char b[4];
        in.read(reinterpret_cast<char*>(b), 4);
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    } else {
        throw std::runtime_error("Unsupported .npy version: " + name);
    }
    std::string h(hlen, '\0');
    in.read(&h[0], hlen);
    if (!in) throw std::runtime_error("Truncated .npy header: " + name);
    NpyHeader out;
    out.data_offset = (major == 1 ? 10 : 12) + hlen;
    auto value_of = [&](const std::string& key) {
        const auto k = h.find("'" + key + "'");
        if (k == std::string::npos) throw std::runtime_error("Malformed .npy header: " + name);
        return h.substr(h.find(':', k) + 1);
    };
    std::string d = strip(value_of("descr"));
    const auto q0 = d.find('\''), q1 = d.find('\'', q0 + 1);
    out.descr = d.substr(q0 + 1, q1 - q0 - 1);
    out.fortran = strip(value_of("fortran_order")).rfind("True", 0) == 0;
    std::string s = value_of("shape");
    s = s.substr(s.find('(') + 1, s.find(')') - s.find('(') - 1);
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ',')) {
        item = strip(item);
        if (!item.empty()) out.shape.push_back(std::stoull(item));
    }
    return out;
}

// ---- .npz writer ------------------------------------------------------------

uint32_t crc32_update(uint32_t crc, const unsigned char* p, size_t n)
{
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        ready = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
    return ~crc;
}

void put16(std::string& b, uint16_t v) { b.push_back((char)(v & 0xFF)); b.push_back((char)(v >> 8)); }
void put32(std::string& b, uint32_t v) { put16(b, (uint16_t)(v & 0xFFFF)); put16(b, (uint16_t)(v >> 16)); }

std::string npy_bytes(const char* descr, const void* data, size_t elem, const std::vector<uint64_t>& shape)
{
    std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (";
    uint64_t count = 1;
    for (size_t i = 0; i < shape.size(); ++i) {
        dict += std::to_string(shape[i]) + (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
        count *= shape[i];
    }
    dict += "), }";
    size_t total = 10 + dict.size() + 1;
    const size_t pad = (64 - total % 64) % 64;
    dict += std::string(pad, ' ') + "\n";
    std::string b("\x93NUMPY\x01\x00", 8);
    put16(b, (uint16_t)dict.size());
    b += dict;
    b.append(static_cast<const char*>(data), count * elem);
    return b;
}

class NpzWriter {
public:
    explicit NpzWriter(const std::string& path) : path_(path), out_(path, std::ios::binary)
    {
        if (!out_) throw std::runtime_error("Failed to open output archive: " + path);
    }
    void add(const std::string& name, const char* descr, const void* data, size_t elem,
             const std::vector<uint64_t>& shape)
    {
        const std::string payload = npy_bytes(descr, data, elem, shape);
        if (payload.size() >= 0xFFFFFFFFull || offset_ >= 0xFFFFFFFFull)
            throw std::runtime_error("archive member exceeds the non-ZIP64 limit: " + name);
        const std::string fname = name + ".npy";
        const uint32_t crc = crc32_update(0, reinterpret_cast<const unsigned char*>(payload.data()),
                                          payload.size());
        std::string lh;
        put32(lh, 0x04034b50u);
        put16(lh, 20); put16(lh, 0); put16(lh, 0);      // version, flags, stored
        put16(lh, 0); put16(lh, 0x21);                   // time, date (1980-01-01)
        put32(lh, crc); put32(lh, (uint32_t)payload.size()); put32(lh, (uint32_t)payload.size());
        put16(lh, (uint16_t)fname.size()); put16(lh, 0);
        lh += fname;
        std::string ch;
        put32(ch, 0x02014b50u);
        put16(ch, 20); put16(ch, 20); put16(ch, 0); put16(ch, 0);
        put16(ch, 0); put16(ch, 0x21);
        put32(ch, crc); put32(ch, (uint32_t)payload.size()); put32(ch, (uint32_t)payload.size());
        put16(ch, (uint16_t)fname.size()); put16(ch, 0); put16(ch, 0); put16(ch, 0); put16(ch, 0);
        put32(ch, 0); put32(ch, (uint32_t)offset_);
        ch += fname;
        central_ += ch;
        out_.write(lh.data(), (std::streamsize)lh.size());
        out_.write(payload.data(), (std::streamsize)payload.size());
        offset_ += lh.size() + payload.size();
        ++count_;
    }
    void close()
    {
        std::string end;
        put32(end, 0x06054b50u);
        put16(end, 0); put16(end, 0);
        put16(end, (uint16_t)count_); put16(end, (uint16_t)count_);
        put32(end, (uint32_t)central_.size()); put32(end, (uint32_t)offset_);
        put16(end, 0);
        out_.write(central_.data(), (std::streamsize)central_.size());
        out_.write(end.data(), (std::streamsize)end.size());
        out_.close();
        if (!out_) throw std::runtime_error("Failed to write output archive: " + path_);
    }

private:
    std::string path_;
    std::ofstream out_;
    std::string central_;
    uint64_t offset_ = 0;
    int count_ = 0;
};

void make_dirs(const std::string& dir)
{
    if (dir.empty()) return;
    std::string cur;
    std::stringstream ss(dir);
    std::string part;
    if (dir[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        if (mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
            throw std::runtime_error("Failed to create directory " + cur);
    }
}

// write_output.hpp:66-86: write to "<path>.tmp", then rename over <path>
template <class F>
void publish(const std::string& path, F&& contents)
{
    const auto slash = path.find_last_of('/');
    if (slash != std::string::npos) make_dirs(path.substr(0, slash));
    const std::string tmp = path + ".tmp";
    try {
        NpzWriter w(tmp);
        contents(w);
        w.close();
    } catch (...) {
        std::remove(tmp.c_str());
        throw;
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        throw std::runtime_error("Failed to publish output archive " + path + ": " +
                                 std::strerror(errno));
    }
}

bool is_dir(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}
bool exists(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}
bool is_file(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

}  // namespace

// The output.variables selection of the step archives, for the driver
// (grt_api.cpp); NULL gives the default selection.  On a bad selection the
// message goes to grt_last_error(NULL).
int grt_output_selection(const char* spec, std::vector<std::string>* names)
{
    try {
        *names = parse_selection(spec ? spec : "id,frequency,x0,x1,x2,x3,k0,k1,k2,k3,I,Q,U,V,phase");
        return GRT_OK;
    } catch (const std::exception& e) {
        grt_set_global_error(e.what());
        return GRT_E_INVALID;
    }
}

extern "C" {

int grt_load_config(const char* path, grt_params* p, grt_run_config* run)
{
    if (!path || !p || !run) {
        grt_set_global_error("path, params and run must not be NULL");
        return GRT_E_INVALID;
    }
    try {
        const RawConfig cfg = parse_flat_toml(path);
        grt_params_default(p);
        // utils.cpp:4-37 and the option structs' defaults
        double M = 1.0, a = 1.0;
        double cam_theta = kPi / 4.0, cam_phi = kPi / 4.0, w = 20.0, h = 20.0, dist = 10000.0;
        uint64_t nphotons = 20;
        int max_steps = 10000;
        double term_pct = 0.9, dlambda = 0.1, atol = 1e-8, rtol = 1e-8;
        uint64_t interval = 1, stride = 1;
        std::string name = "output", numpy_dir = ".", out_dir = "./output";
        std::string variables = "id,frequency,x0,x1,x2,x3,k0,k1,k2,k3,I,Q,U,V,phase";
        bool vacuum = false, scatter = false, absorb = false, emit = false;
        double albedo = 1.0, kabs = 0.0, kemit = 0.0;
        uint64_t seed = 12345u;
        uint64_t nx = 256, ny = 256, bins = 128;
        double smin = 1.0e8, smax = 1.0e22;

        // toml_config.hpp:137-234, in the same order (a failing key reports first)
        const std::string mode = get_string(cfg, "simulation.mode", "image");
        if (mode == "image") p->mode = GRT_MODE_IMAGE;
        else if (mode == "disk") p->mode = GRT_MODE_DISK;
        else throw std::runtime_error("simulation.mode must be 'image' or 'disk'");
        name = get_string(cfg, "simulation.name", name);
        numpy_dir = get_string(cfg, "input.numpy_dir", numpy_dir);
        vacuum = get_bool(cfg, "input.vacuum", vacuum);
        scatter = get_bool(cfg, "scattering.enabled", scatter);
        albedo = get_real(cfg, "scattering.albedo", albedo);
        seed = get_unsigned(cfg, "scattering.seed", seed);
        absorb = get_bool(cfg, "radiative_transfer.absorption_enabled", absorb);
        emit = get_bool(cfg, "radiative_transfer.emission_enabled", emit);
        kabs = get_real(cfg, "radiative_transfer.absorption_coefficient", kabs);
        kemit = get_real(cfg, "radiative_transfer.emission_coefficient", kemit);
        const std::string integ = get_string(cfg, "integration.integrator", "rk45");
        if (integ == "rk4") p->integrator = GRT_RK4;
        else if (integ == "rk45") p->integrator = GRT_RK45;
        else throw std::runtime_error("integration.integrator must be \"rk4\" or \"rk45\"");
        M = get_real(cfg, "black_hole.mass", M);
        a = get_real(cfg, "black_hole.spin", a);
        if (M <= 0.0)
            throw std::runtime_error("black_hole.mass must be positive (got " + std::to_string(M) + ")");
        if (std::fabs(a) > M)
            throw std::runtime_error(
                "black_hole.spin must satisfy |spin| <= mass, i.e. dimensionless spin a/M in "
                "[-1, 1] (got spin=" + std::to_string(a) + ", mass=" + std::to_string(M) +
                "); |spin| > mass is a naked singularity.");
        const double horizon = M + std::sqrt(std::fmax(0.0, M * M - a * a));
        cam_theta = get_real(cfg, "camera.theta", cam_theta);
        cam_phi = get_real(cfg, "camera.phi", cam_phi);
        w = get_real(cfg, "camera.width", w);
        h = get_real(cfg, "camera.height", h);
        dist = get_real(cfg, "camera.distance", dist);
        nphotons = get_unsigned(cfg, "photons.count", nphotons);
        const std::string gen = get_string(cfg, "photons.generator", "blackbody");
        if (gen == "blackbody") p->generator = GRT_GEN_BLACKBODY;
        else if (gen == "power-law" || gen == "power_law" || gen == "powerlaw")
            p->generator = GRT_GEN_POWER_LAW;
        else if (gen == "custom") p->generator = GRT_GEN_CUSTOM;
        else
            throw std::runtime_error("photons.generator must be 'blackbody', 'power-law', or "
                                     "'custom' (got '" + gen + "')");
        p->superphoton_count_normalization =
            get_real(cfg, "photons.superphoton_count_normalization", 1.0);
        p->energy_per_cell_erg = get_real(cfg, "photons.energy_per_cell_erg", 1.0);
        p->power_law_slope = get_real(cfg, "photons.power_law_slope", 2.0);
        p->nu_min_hz = get_real(cfg, "photons.nu_min_hz", 1.0e9);
        p->nu_max_hz = get_real(cfg, "photons.nu_max_hz", 1.0e20);
        p->custom_frequency_hz = get_real(cfg, "photons.custom_frequency_hz", 2.30e11);
        p->length_cm_per_code = get_real(cfg, "units.length_cm_per_code", 1.0);
        p->time_s_per_code = get_real(cfg, "units.time_s_per_code", 1.0);
        p->density_g_cm3_per_code = get_real(cfg, "units.density_g_cm3_per_code", 1.0);
        p->temperature_k_per_code = get_real(cfg, "units.temperature_k_per_code", 1.0);
        p->four_velocity_cm_s_per_code = get_real(cfg, "units.four_velocity_cm_s_per_code", 1.0);
        p->magnetic_gauss_per_code = get_real(cfg, "units.magnetic_gauss_per_code", 1.0);
        const std::string decomposition = get_string(cfg, "parallel.decomposition", "[1, 1, -1]");
        max_steps = get_int(cfg, "integration.max_steps", max_steps);
        term_pct = get_real(cfg, "integration.termination_percent", term_pct);
        dlambda = get_real(cfg, "integration.initial_step", dlambda);
        atol = get_real(cfg, "integration.atol", atol);
        rtol = get_real(cfg, "integration.rtol", rtol);
        const double rmin = get_real(cfg, "termination.r_min", horizon);
        const double rmax = get_real(cfg, "termination.r_max", 1.5 * dist);
        interval = get_unsigned(cfg, "output.interval", interval);
        stride = get_unsigned(cfg, "output.stride", stride);
        out_dir = get_string(cfg, "output.directory", out_dir);
        variables = get_string(cfg, "output.variables", variables);
        nx = get_unsigned(cfg, "output.image_nx", nx);
        ny = get_unsigned(cfg, "output.image_ny", ny);
        bins = get_unsigned(cfg, "output.spectrum_bins", bins);
        smin = get_real(cfg, "output.spectrum_min_hz", smin);
        smax = get_real(cfg, "output.spectrum_max_hz", smax);

        // toml_config.hpp:236-277
        if (nphotons == 0 || nphotons > 100000000000ull)
            throw std::runtime_error("photons.count must be in [1, 100000000000]");
        if (max_steps <= 0) throw std::runtime_error("integration.max_steps must be positive");
        if (interval == 0) throw std::runtime_error("output.interval must be positive");
        if (stride == 0) throw std::runtime_error("output.stride must be positive");
        if (numpy_dir.empty()) throw std::runtime_error("input.numpy_dir must not be empty");
        if (out_dir.empty()) throw std::runtime_error("output.directory must not be empty");
        if (name.empty()) throw std::runtime_error("simulation.name must not be empty");
        if (nx == 0 || ny == 0 || bins == 0)
            throw std::runtime_error("output image dimensions and spectrum_bins must be positive");
        if (!(smin > 0.0) || !(smax > smin))
            throw std::runtime_error("output requires 0 < spectrum_min_hz < spectrum_max_hz");
        if (!(dist > 0.0) || !(w > 0.0) || !(h > 0.0))
            throw std::runtime_error("camera distance, width, and height must be positive");
        if (!std::isfinite(albedo) || albedo < 0.0 || albedo > 1.0)
            throw std::runtime_error("scattering.albedo must be finite and in [0, 1]");
        if (!std::isfinite(kabs) || kabs < 0.0)
            throw std::runtime_error(
                "radiative_transfer.absorption_coefficient must be finite and non-negative");
        if (!std::isfinite(kemit) || kemit < 0.0)
            throw std::runtime_error(
                "radiative_transfer.emission_coefficient must be finite and non-negative");
        if (term_pct <= 0.0 || term_pct > 1.0)
            throw std::runtime_error("integration.termination_percent must be in (0, 1]");
        const std::vector<std::string> sel = parse_selection(variables);

        p->a_bh = a;
        p->m_bh = M;
        p->r_term_min = rmin;
        p->r_term_max = rmax;
        p->atol = atol;
        p->rtol = rtol;
        p->absorption_enabled = absorb;
        p->emission_enabled = emit;
        p->absorption_coefficient = kabs;
        p->emission_coefficient = kemit;
        p->scattering_enabled = scatter;
        p->albedo = albedo;
        p->seed = seed;
        p->camera_distance = dist;
        p->camera_theta = cam_theta;
        p->camera_phi = cam_phi;
        p->camera_width = w;
        p->camera_height = h;
        std::memset(run, 0, sizeof *run);
        copy_string(run->name, sizeof run->name, name, "simulation.name");
        copy_string(run->numpy_dir, sizeof run->numpy_dir, numpy_dir, "input.numpy_dir");
        copy_string(run->output_dir, sizeof run->output_dir, out_dir, "output.directory");
        copy_string(run->output_variables, sizeof run->output_variables, join(sel),
                    "output.variables");
        copy_string(run->decomposition, sizeof run->decomposition, decomposition,
                    "parallel.decomposition");
        run->vacuum = vacuum ? 1 : 0;
        run->photon_count = nphotons;
        run->max_steps = max_steps;
        run->termination_percent = term_pct;
        run->initial_step = dlambda;
        run->output_interval = interval;
        run->output_stride = stride;
        run->image_nx = nx;
        run->image_ny = ny;
        run->spectrum_bins = bins;
        run->spectrum_min_hz = smin;
        run->spectrum_max_hz = smax;
        // the remaining checks of the physics parameters (photon_generation.hpp
        // :52-75, utils.hpp UnitConversions::validate) are grt_create's
        return GRT_OK;
    } catch (const std::invalid_argument& e) {
        grt_set_global_error(std::string("invalid number in config: ") + e.what());
    } catch (const std::out_of_range& e) {
        grt_set_global_error(std::string("number out of range in config: ") + e.what());
    } catch (const std::exception& e) {
        grt_set_global_error(e.what());
    }
    return GRT_E_INVALID;
}

int grt_read_npy(const char* path, double* dst, uint64_t capacity, uint64_t shape[4],
                 int32_t* ndim)
{
    if (!path || !shape || !ndim) {
        grt_set_global_error("path, shape and ndim must not be NULL");
        return GRT_E_INVALID;
    }
    try {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::runtime_error(std::string("Failed to open numpy file: ") + path);
        const NpyHeader h = read_npy_header(in, path);
        if (h.descr != "<f8" && h.descr != "=f8")
            throw std::runtime_error(
                std::string("Grid arrays must use native/little-endian float64 values: ") + path);
        if (h.fortran)
            throw std::runtime_error(std::string("Grid arrays must be C-ordered: ") + path);
        if (h.shape.size() > 4)
            throw std::runtime_error(std::string("more than four dimensions: ") + path);
        uint64_t count = 1;
        for (size_t i = 0; i < 4; ++i) shape[i] = i < h.shape.size() ? h.shape[i] : 0;
        for (uint64_t s : h.shape) count *= s;
        *ndim = (int32_t)h.shape.size();
        if (!dst) return GRT_OK;
        if (count > capacity) throw std::runtime_error(std::string("destination too small: ") + path);
        in.seekg((std::streamoff)h.data_offset);
        in.read(reinterpret_cast<char*>(dst), (std::streamsize)(count * sizeof(double)));
        if (!in)
            throw std::runtime_error(std::string("Unexpected EOF while reading numpy data: ") + path);
        return GRT_OK;
    } catch (const std::exception& e) {
        grt_set_global_error(e.what());
        return GRT_E_INVALID;
    }
}

int grt_load_numpy_fields(grt_ctx* ctx, const char* directory)
{
    if (!directory) return grt_set_error(ctx, GRT_E_INVALID, "directory must not be NULL");
    try {
        const std::string root = directory;
        if (!exists(root)) throw std::runtime_error("NumPy directory does not exist: " + root);
        if (!is_dir(root)) throw std::runtime_error("NumPy input path is not a directory: " + root);
        const char* names[7] = {"r.npy", "theta.npy", "phi.npy", "rho.npy", "Tgas.npy", "vel.npy",
                                "mag.npy"};
        std::string missing;
        for (const char* n : names) {
            if (!is_file(root + "/" + n)) missing += (missing.empty() ? "" : ", ") + std::string(n);
        }
        if (!missing.empty())
            throw std::runtime_error("NumPy directory '" + root +
                                     "' is missing required file(s): " + missing);
        uint64_t shape[4];
        int32_t nd = 0;
        const std::string rpath = root + "/r.npy";
        if (grt_read_npy(rpath.c_str(), nullptr, 0, shape, &nd) != GRT_OK)
            throw std::runtime_error(grt_last_error(nullptr));
        if (nd != 3 || !shape[0] || !shape[1] || !shape[2])
            throw std::runtime_error("Coordinate grid must have shape (nr, ntheta, nphi): " + rpath);
        const uint64_t G = shape[0] * shape[1] * shape[2];
        std::vector<std::vector<double>> a(7);
        for (int f = 0; f < 7; ++f) {
            const std::string fp = root + "/" + names[f];
            const bool vec4 = f >= 5;
            uint64_t sh[4];
            int32_t d = 0;
            if (grt_read_npy(fp.c_str(), nullptr, 0, sh, &d) != GRT_OK)
                throw std::runtime_error(grt_last_error(nullptr));
            if (d != (vec4 ? 4 : 3) || sh[0] != shape[0] || sh[1] != shape[1] || sh[2] != shape[2])
                throw std::runtime_error("Grid field has incompatible dimensions: " + fp);
            if (vec4 && sh[3] != 4)
                throw std::runtime_error("Four-vector field must have four components: " + fp);
            a[f].resize(vec4 ? 4 * G : G);
            if (grt_read_npy(fp.c_str(), a[f].data(), a[f].size(), sh, &d) != GRT_OK)
                throw std::runtime_error(grt_last_error(nullptr));
        }
        if (!ctx) return GRT_OK;   // check only
        return grt_set_grid(ctx, (int32_t)shape[0], (int32_t)shape[1], (int32_t)shape[2],
                            a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(),
                            a[5].data(), a[6].data());
    } catch (const std::exception& e) {
        return grt_set_error(ctx, GRT_E_INVALID, e.what());
    }
}

int grt_write_image_products_npz(const char* path, int64_t nx, int64_t ny, int64_t bins,
                                 const double* image, const double* spectrum_nu,
                                 const double* spectrum_I, uint64_t arrived)
{
    if (!path || !image || !spectrum_nu || !spectrum_I || nx <= 0 || ny <= 0 || bins <= 0) {
        grt_set_global_error("invalid arguments to grt_write_image_products_npz");
        return GRT_E_INVALID;
    }
    try {
        const size_t np = (size_t)nx * (size_t)ny;
        publish(path, [&](NpzWriter& w) {
            const std::vector<uint64_t> sh2{(uint64_t)ny, (uint64_t)nx};
            w.add("image_I", "<f8", image, 8, sh2);
            w.add("image_Q", "<f8", image + np, 8, sh2);
            w.add("image_U", "<f8", image + 2 * np, 8, sh2);
            w.add("image_V", "<f8", image + 3 * np, 8, sh2);
            w.add("spectrum_frequency_hz", "<f8", spectrum_nu, 8, {(uint64_t)bins});
            w.add("spectrum_I", "<f8", spectrum_I, 8, {(uint64_t)bins});
            w.add("arrived_count", "<u8", &arrived, 8, {1});
        });
        return GRT_OK;
    } catch (const std::exception& e) {
        grt_set_global_error(e.what());
        return GRT_E_INVALID;
    }
}

int grt_write_photons_npz(const char* path, const grt_photons* ph, uint64_t n,
                          const char* variables, uint64_t stride)
{
    if (!path || !ph || stride == 0) {
        grt_set_global_error("invalid arguments to grt_write_photons_npz");
        return GRT_E_INVALID;
    }
    try {
        const std::vector<std::string> sel =
            parse_selection(variables ? variables : "id,frequency,x0,x1,x2,x3,k0,k1,k2,k3,I,Q,U,V,phase");
        const uint64_t m = n == 0 ? 0 : 1 + (n - 1) / stride;
        publish(path, [&](NpzWriter& w) {
            for (const std::string& name : sel) {
                const void* src = nullptr;
                size_t elem = 8;
                const char* descr = "<f8";
                if (name == "id") { src = ph->id; descr = "<u8"; }
                else if (name == "frequency") src = ph->frequency;
                else if (name == "emission_frame_energy") src = ph->emission_frame_energy;
                else if (name == "x0") src = ph->x0;
                else if (name == "x1") src = ph->x1;
                else if (name == "x2") src = ph->x2;
                else if (name == "x3") src = ph->x3;
                else if (name == "k0") src = ph->k0;
                else if (name == "k1") src = ph->k1;
                else if (name == "k2") src = ph->k2;
                else if (name == "k3") src = ph->k3;
                else if (name == "I") src = ph->I;
                else if (name == "Q") src = ph->Q;
                else if (name == "U") src = ph->U;
                else if (name == "V") src = ph->V;
                else if (name == "dlambda") src = ph->dlambda;
                else if (name == "terminate") { src = ph->terminate; elem = 1; descr = "|u1"; }
                else if (name == "phase") { src = ph->phase; elem = 1; descr = "|u1"; }
                else if (name == "theta_disp") src = ph->theta_disp;
                else if (name == "phi_disp") src = ph->phi_disp;
                if (!src) throw std::runtime_error("photon array " + name + " is NULL");
                std::vector<char> buf(m * elem);
                for (uint64_t i = 0; i < m; ++i)
                    std::memcpy(&buf[i * elem], static_cast<const char*>(src) + i * stride * elem, elem);
                w.add(name, descr, buf.data(), elem, {m});
            }
        });
        return GRT_OK;
    } catch (const std::exception& e) {
        grt_set_global_error(e.what());
        return GRT_E_INVALID;
    }
}

}  // extern "C"
