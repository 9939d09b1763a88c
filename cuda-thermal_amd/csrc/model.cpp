// model.cpp - host model builder and case-folder reader.
//
// Conductances follow the finite-volume rule of src/stack3d.h:9-25: two
// half-cells in series across every face, G = area / (d1/(2 k1) + d2/(2 k2)),
// evaluated in double and stored in float, with the same operation order as
// stack3d.h:112-148 so that the float arrays are bit-identical to the
// reference's for the same inputs (tests/test_model.py checks them against
// fixtures written by the reference's own build).
#include "model.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace th {

static thread_local std::string g_err;
static thread_local int g_code = 0;

int fail(int code, const std::string& msg) {
    g_code = code;
    g_err = msg;
    return code;
}
void clear_error() {
    g_code = 0;
    g_err.clear();
}
const char* last_error() { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Geometry and materials
// ---------------------------------------------------------------------------

int make_model(const std::vector<Layer>& layers, int nx, int ny, double Lx, double Ly,
               double h_top, double h_bot, Model& m) {
    if (nx < 1 || ny < 1 || layers.empty() || !(Lx > 0) || !(Ly > 0))
        return fail(TH_ERR_ARG, "model: need nx, ny >= 1, at least one layer and a positive die size");
    m = Model();
    m.layers = layers;
    m.nx = nx;
    m.ny = ny;
    m.Lx = Lx;
    m.Ly = Ly;
    m.dx = Lx / nx;
    m.dy = Ly / ny;
    m.h_top = h_top;
    m.h_bot = h_bot;
    for (size_t L = 0; L < layers.size(); ++L) {
        if (layers[L].planes < 1) return fail(TH_ERR_ARG, "model: every layer needs at least one plane");
        const double t = layers[L].thickness / layers[L].planes;
        for (int c = 0; c < layers[L].planes; ++c) {
            m.plane_layer.push_back((int)L);
            m.dz.push_back(t);
        }
    }
    m.nz = (int)m.dz.size();
    if (m.nz > TH_MAXZ) {
        char buf[96];
        snprintf(buf, sizeof buf, "Too many z-planes (%d > %d)", m.nz, TH_MAXZ);
        return fail(TH_ERR_ZLIMIT, buf);
    }
    m.tsv_frac.assign(m.ncells(), 0.0);
    m.P.assign(m.ncells(), 0.0);
    return TH_OK;
}

// The reference's two-tier stack, bottom to top (stack3d.h:153-162).
std::vector<Layer> default_layers() {
    return {
        {"Si tier 2 (bottom)", 50e-6, 130.0, 1.66e6, 2, true},
        {"bond + BEOL", 10e-6, 2.0, 1.70e6, 1, false},
        {"Si tier 1 (top)", 50e-6, 130.0, 1.66e6, 2, true},
        {"TIM", 25e-6, 5.0, 2.00e6, 1, false},
        {"Cu spreader", 1e-3, 400.0, 3.45e6, 4, false},
    };
}

// Conductance of two half-cells in series: area / (d1/(2 k1) + d2/(2 k2)).
static inline double series(double area, double d1, double k1, double d2, double k2) {
    return area / (d1 / (2 * k1) + d2 / (2 * k2));
}

void finalize(Model& m) {
    const size_t n = m.ncells(), plane = m.nxy();
    const double area_z = m.dx * m.dy;
    std::vector<double> kl(n), kv(n);  // lateral and vertical conductivity per cell
    m.cap.assign(n, 0.0f);
    for (int l = 0; l < m.nz; ++l) {
        const Layer& L = m.layers[m.plane_layer[l]];
        for (size_t c = 0; c < plane; ++c) {
            const size_t id = (size_t)l * plane + c;
            const double f = m.tsv_frac[id];
            kl[id] = kxy_mix(L.k, m.tsv_k, f);
            kv[id] = kz_mix(L.k, m.tsv_k, f);
            m.cap[id] = (float)((f * m.tsv_rho_c + (1 - f) * L.rho_c) * area_z * m.dz[l]);
        }
    }
    m.gE.assign(n, 0.0f);
    m.gN.assign(n, 0.0f);
    m.gU.assign(n, 0.0f);
    m.gB.assign(plane, 0.0f);
    for (int l = 0; l < m.nz; ++l) {
        const double h = m.dz[l];
        const bool top = (l == m.nz - 1);
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) {
                const size_t id = m.at(i, j, l);
                if (i + 1 < m.nx) m.gE[id] = (float)series(m.dy * h, m.dx, kl[id], m.dx, kl[id + 1]);
                if (j + 1 < m.ny) m.gN[id] = (float)series(m.dx * h, m.dy, kl[id], m.dy, kl[id + m.nx]);
                // Up: to the plane above, or through the heat sink on the top plane.
                const double r_up = top ? h / (2 * kv[id]) + 1.0 / m.h_top
                                        : h / (2 * kv[id]) + m.dz[l + 1] / (2 * kv[id + plane]);
                m.gU[id] = (float)(area_z / r_up);
                if (l == 0 && m.h_bot > 0) m.gB[id] = (float)(area_z / (h / (2 * kv[id]) + 1.0 / m.h_bot));
            }
    }
    m.finalized = true;
}

// Calls fn(in-plane cell, covered fraction) for each cell the rectangle
// [x0, x0+w] x [y0, y0+h] covers by more than a rounding sliver
// (stack3d.h:177-191).
template <class Fn>
static void covered_cells(const Model& m, double x0, double y0, double w, double h, Fn fn) {
    auto span = [](double a0, double a1, double b0, double b1) {
        return std::max(0.0, std::min(a1, b1) - std::max(a0, b0));
    };
    const int i0 = std::max(0, (int)std::floor(x0 / m.dx));
    const int i1 = std::min(m.nx - 1, (int)std::ceil((x0 + w) / m.dx));
    const int j0 = std::max(0, (int)std::floor(y0 / m.dy));
    const int j1 = std::min(m.ny - 1, (int)std::ceil((y0 + h) / m.dy));
    const double cell = m.dx * m.dy;
    for (int j = j0; j <= j1; ++j)
        for (int i = i0; i <= i1; ++i) {
            const double a = span(i * m.dx, (i + 1) * m.dx, x0, x0 + w) * span(j * m.dy, (j + 1) * m.dy, y0, y0 + h);
            if (a > 1e-9 * cell) fn((size_t)j * m.nx + i, a / cell);
        }
}

void power_block(Model& m, int layer, double x0, double y0, double w, double h, double watts) {
    const double per_area = watts / (w * h);
    const int planes = m.layers[layer].planes;
    for (int l = 0; l < m.nz; ++l) {
        if (m.plane_layer[l] != layer) continue;
        double* P = m.P.data() + (size_t)l * m.nxy();
        covered_cells(m, x0, y0, w, h, [&](size_t c, double frac) { P[c] += per_area * frac * m.dx * m.dy / planes; });
    }
}

// Max and area-weighted mean of T over a rectangle of a layer's planes
// (main3d.cu:286-301).
void block_temps(const Model& m, const double* T, int layer, double x0, double y0, double w, double h, double& tmax,
                 double& tavg) {
    tmax = -1e30;
    double sum = 0, wsum = 0;
    for (int l = 0; l < m.nz; ++l) {
        if (m.plane_layer[l] != layer) continue;
        covered_cells(m, x0, y0, w, h, [&](size_t c, double frac) {
            const double t = T[(size_t)l * m.nxy() + c];
            tmax = std::max(tmax, t);
            sum += frac * t;
            wsum += frac;
        });
    }
    tavg = wsum > 0 ? sum / wsum : 0.0;
}

void tsv_region(Model& m, int layer, double x0, double y0, double w, double h, double f) {
    for (int l = 0; l < m.nz; ++l) {
        if (m.plane_layer[l] != layer) continue;
        double* t = m.tsv_frac.data() + (size_t)l * m.nxy();
        covered_cells(m, x0, y0, w, h, [&](size_t c, double frac) { t[c] = std::min(1.0, t[c] + f * frac); });
    }
    m.finalized = false;
}

// Cells whose centres fall inside the square share the power equally
// (stack3d.h:230-244).
bool hotspot(Model& m, int layer, double xc, double yc, double size, double watts) {
    std::vector<size_t> hit;
    for (int l = 0; l < m.nz; ++l) {
        if (m.plane_layer[l] != layer) continue;
        for (int j = 0; j < m.ny; ++j)
            for (int i = 0; i < m.nx; ++i) {
                const double x = (i + 0.5) * m.dx, y = (j + 0.5) * m.dy;
                if (std::fabs(x - xc) <= size / 2 && std::fabs(y - yc) <= size / 2) hit.push_back(m.at(i, j, l));
            }
    }
    if (hit.empty()) return false;
    for (size_t c : hit) m.P[c] += watts / (double)hit.size();
    return true;
}

double total_power(const Model& m) {
    double s = 0;
    for (double p : m.P) s += p;
    return s;
}

// Exact temperature of each plane for laterally uniform power with an
// adiabatic bottom and no TSVs: a 1D resistor chain, from the double layer
// data (stack3d.h:279-298).
std::vector<double> chain_1d(const Model& m) {
    const double area = m.dx * m.dy;
    std::vector<double> g(m.nz), below(m.nz), T(m.nz);
    for (int l = 0; l < m.nz; ++l) {
        const double k = m.layers[m.plane_layer[l]].k;
        const double r = (l + 1 < m.nz) ? m.dz[l] / (2 * k) + m.dz[l + 1] / (2 * m.layers[m.plane_layer[l + 1]].k)
                                        : m.dz[l] / (2 * k) + 1.0 / m.h_top;
        g[l] = area / r;
    }
    double acc = 0;  // power generated in planes 0..l, which crosses the top face of plane l
    for (int l = 0; l < m.nz; ++l) {
        for (size_t c = 0; c < m.nxy(); ++c) acc += m.P[(size_t)l * m.nxy() + c];
        below[l] = acc;
    }
    const double cols = (double)m.nxy();
    T[m.nz - 1] = below[m.nz - 1] / (g[m.nz - 1] * cols);
    for (int l = m.nz - 2; l >= 0; --l) T[l] = T[l + 1] + below[l] / (g[l] * cols);
    return T;
}

double heat_out(const Model& m, const double* T) {
    double q = 0;
    const size_t top = (size_t)(m.nz - 1) * m.nxy();
    for (size_t c = 0; c < m.nxy(); ++c) {
        q += (double)m.gU[top + c] * T[top + c];
        q += (double)m.gB[c] * T[c];
    }
    return q;
}

// ---------------------------------------------------------------------------
// Case folders: stack.csv, floorplan.csv, optional tsv.csv (case3d.h:1-13)
// ---------------------------------------------------------------------------

static std::string strip(const std::string& s) {
    const char* ws = " \t\r\n\"";
    const size_t a = s.find_first_not_of(ws), b = s.find_last_not_of(ws);
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

static std::vector<std::string> fields(const std::string& line) {
    std::vector<std::string> out;
    std::stringstream ss(line);
    std::string f;
    while (std::getline(ss, f, ',')) out.push_back(strip(f));
    return out;
}

struct ParseError {
    std::string what;
};
static double num(const std::string& s) {
    const char* p = s.c_str();
    char* end = nullptr;
    const double v = std::strtod(p, &end);
    if (end == p) throw ParseError{"'" + s + "'"};
    return v;
}
static int whole(const std::string& s) {
    const char* p = s.c_str();
    char* end = nullptr;
    const long v = std::strtol(p, &end, 10);
    if (end == p) throw ParseError{"'" + s + "'"};
    return (int)v;
}

// Data rows of a CSV file, header skipped; '#' lines go to `settings`.
static bool csv_rows(const std::string& path, std::vector<std::vector<std::string>>& rows,
                     std::vector<std::vector<std::string>>* settings) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    bool header = false;
    while (std::getline(in, line)) {
        const std::string t = strip(line);
        if (t.empty()) continue;
        if (t[0] == '#') {
            if (settings) settings->push_back(fields(strip(t.substr(1))));
            continue;
        }
        if (!header) {
            header = true;
            continue;
        }
        rows.push_back(fields(t));
    }
    return true;
}

static int layer_named(const std::vector<Layer>& layers, const std::string& name) {
    for (size_t L = 0; L < layers.size(); ++L)
        if (layers[L].name == name) return (int)L;
    return -1;
}

bool read_case(const std::string& dir, Case& c, std::string& err) {
    c = Case();
    std::vector<std::vector<std::string>> rows, settings;
    try {
        if (!csv_rows(dir + "/stack.csv", rows, &settings)) {
            err = "cannot open " + dir + "/stack.csv";
            return false;
        }
        for (const auto& s : settings) {
            if (s.empty()) continue;
            if (s[0] == "die_mm" && s.size() >= 3) {
                c.Lx = num(s[1]) * 1e-3;
                c.Ly = num(s[2]) * 1e-3;
            } else if (s[0] == "h_top" && s.size() >= 2) {
                c.h_top = num(s[1]);
            } else if (s[0] == "h_bottom" && s.size() >= 2) {
                c.h_bot = num(s[1]);
            } else if (s[0] == "ambient_C" && s.size() >= 2) {
                c.T_amb = num(s[1]);
            }
        }
        for (const auto& r : rows) {
            if (r.size() < 6) {
                err = "stack.csv: expected 6 columns";
                return false;
            }
            c.layers.push_back({r[0], num(r[1]) * 1e-6, num(r[2]), num(r[3]), whole(r[4]), whole(r[5]) != 0});
        }
        if (c.layers.empty()) {
            err = "stack.csv has no layers";
            return false;
        }
        rows.clear();
        if (!csv_rows(dir + "/floorplan.csv", rows, nullptr)) {
            err = "cannot open " + dir + "/floorplan.csv";
            return false;
        }
        for (const auto& r : rows) {
            if (r.size() < 7) {
                err = "floorplan.csv: expected 7 columns";
                return false;
            }
            const int L = layer_named(c.layers, r[1]);
            if (L < 0) {
                err = "floorplan.csv: unknown layer '" + r[1] + "'";
                return false;
            }
            if (!c.layers[L].active) {
                err = "floorplan.csv: layer '" + r[1] + "' is not active";
                return false;
            }
            Block b{r[0], L, num(r[2]) * 1e-3, num(r[3]) * 1e-3, num(r[4]) * 1e-3, num(r[5]) * 1e-3, num(r[6])};
            if (b.x0 < -1e-12 || b.y0 < -1e-12 || b.x0 + b.w > c.Lx * (1 + 1e-9) || b.y0 + b.h > c.Ly * (1 + 1e-9)) {
                err = "floorplan.csv: block '" + b.name + "' extends outside the die";
                return false;
            }
            c.blocks.push_back(b);
        }
        rows.clear();
        if (csv_rows(dir + "/tsv.csv", rows, nullptr))
            for (const auto& r : rows) {
                if (r.size() < 6) {
                    err = "tsv.csv: expected 6 columns";
                    return false;
                }
                const int L = layer_named(c.layers, r[0]);
                if (L < 0) {
                    err = "tsv.csv: unknown layer '" + r[0] + "'";
                    return false;
                }
                Tsv t{L, num(r[1]) * 1e-3, num(r[2]) * 1e-3, num(r[3]) * 1e-3, num(r[4]) * 1e-3, num(r[5])};
                if (t.fraction < 0 || t.fraction > 1) {
                    err = "tsv.csv: cu_fraction must be between 0 and 1";
                    return false;
                }
                c.tsvs.push_back(t);
            }
    } catch (const ParseError& e) {
        err = "could not parse a number: " + e.what;
        return false;
    }
    return true;
}

// The example stack of cases/example, built in code (case3d.h:170-190): a
// logic die with four cores, cache and I/O under a memory die, with TSV
// fields beside the cores.
Case builtin_example() {
    Case c;
    c.layers = default_layers();
    const int logic = 0, mem = 2, bond = 1;
    const double mm = 1e-3;
    c.blocks = {
        {"core0", logic, 0.5 * mm, 0.5 * mm, 2.0 * mm, 2.0 * mm, 6.0},
        {"core1", logic, 3.0 * mm, 0.5 * mm, 2.0 * mm, 2.0 * mm, 6.0},
        {"core2", logic, 0.5 * mm, 3.0 * mm, 2.0 * mm, 2.0 * mm, 6.0},
        {"core3", logic, 3.0 * mm, 3.0 * mm, 2.0 * mm, 2.0 * mm, 6.0},
        {"l2_cache", logic, 5.5 * mm, 0.5 * mm, 4.0 * mm, 4.5 * mm, 3.0},
        {"noc_io", logic, 0.5 * mm, 5.5 * mm, 9.0 * mm, 4.0 * mm, 3.0},
        {"dram_banks", mem, 0.0 * mm, 0.0 * mm, 10.0 * mm, 10.0 * mm, 12.0},
        {"dram_ctrl", mem, 4.0 * mm, 4.0 * mm, 2.0 * mm, 2.0 * mm, 3.0},
    };
    for (int L : {bond, mem}) {
        c.tsvs.push_back({L, 2.55 * mm, 0.5 * mm, 0.40 * mm, 4.5 * mm, 0.15});
        c.tsvs.push_back({L, 0.5 * mm, 2.55 * mm, 4.5 * mm, 0.40 * mm, 0.15});
    }
    return c;
}

int case_model(const Case& c, int nx, int ny, bool with_tsv, Model& m) {
    const int rc = make_model(c.layers, nx, ny, c.Lx, c.Ly, c.h_top, c.h_bot, m);
    if (rc) return rc;
    m.T_amb = c.T_amb;
    for (const Block& b : c.blocks) power_block(m, b.layer, b.x0, b.y0, b.w, b.h, b.watts);
    if (with_tsv)
        for (const Tsv& t : c.tsvs) tsv_region(m, t.layer, t.x0, t.y0, t.w, t.h, t.fraction);
    finalize(m);
    return TH_OK;
}

}  // namespace th

// ===========================================================================
// C ABI: model
// ===========================================================================
using namespace th;

extern "C" {

const char* th_last_error(void) { return th::last_error(); }
const char* th_version(void) { return "thermal-mi355x 0.1 (gfx950)"; }

static int checked(const th_model* m) {
    if (!m) return fail(TH_ERR_ARG, "null model");
    return TH_OK;
}
static int valid_layer(const th_model* m, int layer) {
    if (checked(m)) return TH_ERR_ARG;
    if (layer < 0 || layer >= (int)m->m.layers.size()) return fail(TH_ERR_ARG, "layer index out of range");
    return TH_OK;
}

int th_model_create(int nlayers, const char* const* names, const double* thickness_m, const double* k,
                    const double* rho_c, const int* planes, const int* active, int nx, int ny, double Lx_m,
                    double Ly_m, double h_top, double h_bot, th_model** out) {
    clear_error();
    if (!out || nlayers < 1 || !thickness_m || !k || !rho_c || !planes || !active)
        return fail(TH_ERR_ARG, "th_model_create: bad arguments");
    std::vector<Layer> layers;
    for (int L = 0; L < nlayers; ++L)
        layers.push_back({names && names[L] ? names[L] : "layer" + std::to_string(L), thickness_m[L], k[L], rho_c[L],
                          planes[L], active[L] != 0});
    th_model* h = new th_model();
    const int rc = make_model(layers, nx, ny, Lx_m, Ly_m, h_top, h_bot, h->m);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return TH_OK;
}

int th_model_create_default(int nx, int ny, th_model** out) {
    clear_error();
    if (!out) return fail(TH_ERR_ARG, "null out");
    th_model* h = new th_model();
    const int rc = make_model(default_layers(), nx, ny, 10e-3, 10e-3, 2e4, 0.0, h->m);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return TH_OK;
}

int th_model_load_case(const char* dir, int nx, int ny, int with_tsv, th_model** out) {
    clear_error();
    if (!dir || !out) return fail(TH_ERR_ARG, "null argument");
    Case c;
    std::string err;
    if (!read_case(dir, c, err)) return fail(TH_ERR_CASE, err);
    th_model* h = new th_model();
    const int rc = case_model(c, nx, ny, with_tsv != 0, h->m);
    if (rc) {
        delete h;
        return rc;
    }
    h->blocks = c.blocks;
    h->tsv_regions = with_tsv ? (int)c.tsvs.size() : 0;
    *out = h;
    return TH_OK;
}

int th_model_example_case(int nx, int ny, int with_tsv, th_model** out) {
    clear_error();
    if (!out) return fail(TH_ERR_ARG, "null out");
    th_model* h = new th_model();
    const Case c = builtin_example();
    const int rc = case_model(c, nx, ny, with_tsv != 0, h->m);
    if (rc) {
        delete h;
        return rc;
    }
    h->blocks = c.blocks;
    h->tsv_regions = with_tsv ? (int)c.tsvs.size() : 0;
    *out = h;
    return TH_OK;
}

int th_model_block_count(const th_model* m) { return m ? (int)m->blocks.size() : -1; }
int th_model_tsv_count(const th_model* m) { return m ? m->tsv_regions : -1; }

int th_model_block_info(const th_model* m, int k, const char** name, int* layer, double* x0, double* y0, double* w,
                        double* h, double* watts) {
    clear_error();
    if (!m || k < 0 || k >= (int)m->blocks.size()) return fail(TH_ERR_ARG, "no such block");
    const Block& b = m->blocks[k];
    if (name) *name = b.name.c_str();
    if (layer) *layer = b.layer;
    if (x0) *x0 = b.x0;
    if (y0) *y0 = b.y0;
    if (w) *w = b.w;
    if (h) *h = b.h;
    if (watts) *watts = b.watts;
    return TH_OK;
}

int th_model_block_temps(const th_model* m, int k, const double* T, double* tmax, double* tavg) {
    clear_error();
    if (!m || !T || k < 0 || k >= (int)m->blocks.size()) return fail(TH_ERR_ARG, "bad argument");
    const Block& b = m->blocks[k];
    double a = 0, v = 0;
    block_temps(m->m, T, b.layer, b.x0, b.y0, b.w, b.h, a, v);
    if (tmax) *tmax = a;
    if (tavg) *tavg = v;
    return TH_OK;
}

void th_model_destroy(th_model* m) { delete m; }

int th_model_add_power_block(th_model* m, int layer, double x0, double y0, double w, double h, double watts) {
    clear_error();
    if (valid_layer(m, layer)) return TH_ERR_ARG;
    if (!(w > 0) || !(h > 0)) return fail(TH_ERR_ARG, "power block needs a positive size");
    power_block(m->m, layer, x0, y0, w, h, watts);
    return TH_OK;
}
int th_model_add_uniform_power(th_model* m, int layer, double watts) {
    clear_error();
    if (valid_layer(m, layer)) return TH_ERR_ARG;
    power_block(m->m, layer, 0, 0, m->m.Lx, m->m.Ly, watts);
    return TH_OK;
}
int th_model_add_hotspot(th_model* m, int layer, double xc, double yc, double size, double watts) {
    clear_error();
    if (valid_layer(m, layer)) return TH_ERR_ARG;
    if (!hotspot(m->m, layer, xc, yc, size, watts)) return fail(TH_ERR_NO_CELLS, "Hotspot covers no cells");
    return TH_OK;
}
int th_model_add_tsv_region(th_model* m, int layer, double x0, double y0, double w, double h, double f) {
    clear_error();
    if (valid_layer(m, layer)) return TH_ERR_ARG;
    if (f < 0 || f > 1) return fail(TH_ERR_ARG, "TSV fraction must be between 0 and 1");
    tsv_region(m->m, layer, x0, y0, w, h, f);
    return TH_OK;
}
int th_model_finalize(th_model* m) {
    clear_error();
    if (checked(m)) return TH_ERR_ARG;
    finalize(m->m);
    return TH_OK;
}
int th_model_set_power(th_model* m, const double* P) {
    clear_error();
    if (checked(m) || !P) return fail(TH_ERR_ARG, "null argument");
    std::copy(P, P + m->m.ncells(), m->m.P.begin());
    return TH_OK;
}
int th_model_dims(const th_model* m, int* nx, int* ny, int* nz, double* dx, double* dy) {
    if (checked(m)) return TH_ERR_ARG;
    if (nx) *nx = m->m.nx;
    if (ny) *ny = m->m.ny;
    if (nz) *nz = m->m.nz;
    if (dx) *dx = m->m.dx;
    if (dy) *dy = m->m.dy;
    return TH_OK;
}
int th_model_layer_count(const th_model* m) { return m ? (int)m->m.layers.size() : -1; }
int th_model_plane_layer(const th_model* m, int* pl) {
    if (checked(m) || !pl) return TH_ERR_ARG;
    std::copy(m->m.plane_layer.begin(), m->m.plane_layer.end(), pl);
    return TH_OK;
}
int th_model_layer_info(const th_model* m, int layer, double* thickness_m, double* k, double* rho_c,
                        int* planes, int* active, const char** name) {
    if (valid_layer(m, layer)) return TH_ERR_ARG;
    const Layer& L = m->m.layers[layer];
    if (thickness_m) *thickness_m = L.thickness;
    if (k) *k = L.k;
    if (rho_c) *rho_c = L.rho_c;
    if (planes) *planes = L.planes;
    if (active) *active = L.active ? 1 : 0;
    if (name) *name = L.name.c_str();
    return TH_OK;
}
int th_model_settings(const th_model* m, double* Lx, double* Ly, double* h_top, double* h_bot, double* T_amb) {
    if (checked(m)) return TH_ERR_ARG;
    if (Lx) *Lx = m->m.Lx;
    if (Ly) *Ly = m->m.Ly;
    if (h_top) *h_top = m->m.h_top;
    if (h_bot) *h_bot = m->m.h_bot;
    if (T_amb) *T_amb = m->m.T_amb;
    return TH_OK;
}
int th_model_get_f32(const th_model* m, const char* which, float* out) {
    clear_error();
    if (checked(m) || !which || !out) return fail(TH_ERR_ARG, "null argument");
    const Model& M = m->m;
    if (!M.finalized) return fail(TH_ERR_NOT_FINALIZED, "Model3D used before finalize()");
    const std::vector<float>* v = !strcmp(which, "gE")    ? &M.gE
                                  : !strcmp(which, "gN")  ? &M.gN
                                  : !strcmp(which, "gU")  ? &M.gU
                                  : !strcmp(which, "gB")  ? &M.gB
                                  : !strcmp(which, "cap") ? &M.cap
                                                          : nullptr;
    if (!v) return fail(TH_ERR_ARG, std::string("unknown float array ") + which);
    std::copy(v->begin(), v->end(), out);
    return TH_OK;
}
int th_model_get_f64(const th_model* m, const char* which, double* out) {
    clear_error();
    if (checked(m) || !which || !out) return fail(TH_ERR_ARG, "null argument");
    const Model& M = m->m;
    const std::vector<double>* v = !strcmp(which, "P")          ? &M.P
                                   : !strcmp(which, "tsv_frac") ? &M.tsv_frac
                                   : !strcmp(which, "dz")       ? &M.dz
                                                                : nullptr;
    if (!v) return fail(TH_ERR_ARG, std::string("unknown double array ") + which);
    std::copy(v->begin(), v->end(), out);
    return TH_OK;
}
double th_model_total_power(const th_model* m) { return m ? total_power(m->m) : 0.0; }
int th_model_analytic_1d(const th_model* m, double* T) {
    if (checked(m) || !T) return TH_ERR_ARG;
    const std::vector<double> t = chain_1d(m->m);
    std::copy(t.begin(), t.end(), T);
    return TH_OK;
}
double th_model_heat_out(const th_model* m, const double* T) {
    if (!m || !T || !m->m.finalized) return 0.0;
    return heat_out(m->m, T);
}
int th_model_describe(const th_model* m, th_model_desc* d) {
    clear_error();
    if (checked(m) || !d) return fail(TH_ERR_ARG, "null argument");
    if (!m->m.finalized) return fail(TH_ERR_NOT_FINALIZED, "Model3D used before finalize()");
    const Model& M = m->m;
    d->nx = M.nx;
    d->ny = M.ny;
    d->nz = M.nz;
    d->dx = M.dx;
    d->dy = M.dy;
    d->gE = M.gE.data();
    d->gN = M.gN.data();
    d->gU = M.gU.data();
    d->gB = M.gB.data();
    d->cap = M.cap.data();
    d->P = M.P.data();
    return TH_OK;
}

}  // extern "C"
