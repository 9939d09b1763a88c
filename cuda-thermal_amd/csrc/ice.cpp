// ice.cpp - 3D-ICE interchange: writes a case folder as a 3D-ICE stack
// description (case.stk, one floorplan per die, TSV layout files) and reads
// 3D-ICE's Tmap / Tflp result files back (src/case3d.h:192-400,
// src/main3d.cu:380-416).
//
// The files are text for an external tool, so their wording, number formats
// and ordering are the reference exporter's (a case exported here and there
// gives the same bytes); the code that produces them is this file's own:
// every file is assembled in memory and written in one go, and the checks
// run before anything is written.
//
// Model mapping (the reference's): every z-plane of an inactive layer is one
// 3D-ICE layer; every active layer (one plane) is a die with one source
// layer carrying its power blocks. Units: um, W/(um K), J/(um^3 K),
// W/(um^2 K), kelvin. A plane with TSVs gets a layout file: one anisotropic
// material per distinct copper fraction (rounded to 1e-9), with the model's
// own effective kxy, kz and rho*c, and its cells as rectangles, one per run of
// equal cells along a row.
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "model.h"

namespace th {
namespace {

// printf into a growing string
void put(std::string& s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void put(std::string& s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    const int n = std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n < (int)sizeof buf) {
        s.append(buf, n > 0 ? (size_t)n : 0);
        return;
    }
    std::vector<char> big((size_t)n + 1);
    va_start(ap, fmt);
    std::vsnprintf(big.data(), big.size(), fmt, ap);
    va_end(ap);
    s.append(big.data(), (size_t)n);
}

bool write_file(const std::string& path, const std::string& text) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    return std::fclose(f) == 0 && ok;
}

// Block names as 3D-ICE identifiers: anything but letters and digits -> '_'.
std::string ident(const std::string& s) {
    std::string out(s);
    for (char& c : out)
        if (!std::isalnum((unsigned char)c)) c = '_';
    return out;
}

double span(double a0, double a1, double b0, double b1) { return std::max(0.0, std::min(a1, b1) - std::max(a0, b0)); }

// Layout file of plane l (layer ly) of the model: the materials by copper
// fraction, then each material's rectangles.
std::string layout_text(const Model& m, int l, const Layer& ly, double cell_um) {
    std::map<long long, std::string> runs;  // fraction key (1e-9 units) -> its rectangles
    const double* f = m.tsv_frac.data() + (size_t)l * m.nxy();
    for (int j = 0; j < m.ny; ++j) {
        int i = 0;
        while (i < m.nx) {
            const long long key = std::llround(f[(size_t)j * m.nx + i] * 1e9);
            int end = i + 1;
            while (end < m.nx && std::llround(f[(size_t)j * m.nx + end] * 1e9) == key) ++end;
            if (key != 0)
                put(runs[key], "   rectangle (%.6f, %.6f, %.6f, %.6f) ;\n", i * cell_um, j * cell_um,
                    (end - i) * cell_um, cell_um);
            i = end;
        }
    }
    std::string s;
    put(s, "// TSV layout for %s (generated by thermal3d export3dice)\n\n", ly.name.c_str());
    int id = 0;
    for (const auto& r : runs) {
        const double fr = r.first * 1e-9;
        const double kxy = kxy_mix(ly.k, m.tsv_k, fr) * 1e-6;
        put(s,
            "material TSV%d :   // copper fraction %.6f\n"
            "   thermal conductivity     %.10e, %.10e, %.10e ;\n"
            "   volumetric heat capacity %.10e ;\n\n",
            id++, fr, kxy, kxy, kz_mix(ly.k, m.tsv_k, fr) * 1e-6, (fr * m.tsv_rho_c + (1 - fr) * ly.rho_c) * 1e-18);
    }
    id = 0;
    for (const auto& r : runs) {
        put(s, "TSV%d :\n", id++);
        s += r.second;
        s += "\n";
    }
    return s;
}

}  // namespace

// The stack description of case c on cells of cell_um micrometres, written
// into outdir (which must exist). nx, ny: the lateral grid used.
int export_3dice(const Case& c, double cell_um, const std::string& outdir, int* nx_out, int* ny_out) {
    if (!(cell_um > 0)) return fail(TH_ERR_ARG, "cell size must divide the die size");
    const double Lx = c.Lx * 1e6, Ly = c.Ly * 1e6;
    const double cx = Lx / cell_um, cy = Ly / cell_um;
    if (std::fabs(cx - std::round(cx)) > 1e-9 || std::fabs(cy - std::round(cy)) > 1e-9)
        return fail(TH_ERR_CASE, "cell size must divide the die size");
    const int nx = (int)std::lround(cx), ny = (int)std::lround(cy);
    Model m;
    const int rc = case_model(c, nx, ny, true, m);
    if (rc) return rc;
    const size_t nl = c.layers.size();
    std::vector<int> first(nl, 0);  // first (lowest) plane of each layer
    for (int l = m.nz - 1; l >= 0; --l) first[m.plane_layer[l]] = l;
    auto has_tsv = [&](int l) {
        const double* f = m.tsv_frac.data() + (size_t)l * m.nxy();
        for (size_t q = 0; q < m.nxy(); ++q)
            if (f[q] > 0) return true;
        return false;
    };
    // The checks, in the reference's order (the first failing one is reported).
    for (size_t L = 0; L < nl; ++L)
        if (c.layers[L].active && has_tsv(first[L]))
            return fail(TH_ERR_CASE, "TSVs in active layer '" + c.layers[L].name +
                                         "' can't be exported (3D-ICE has no material layouts in source layers)");
    if (c.h_bot > 0) return fail(TH_ERR_CASE, "a bottom heat sink is not exported");
    for (size_t L = 0; L < nl; ++L)
        if (c.layers[L].active && c.layers[L].planes != 1)
            return fail(TH_ERR_CASE, "active layer '" + c.layers[L].name + "' must have 1 plane for 3D-ICE");
    for (size_t a = 0; a < c.blocks.size(); ++a)
        for (size_t b = a + 1; b < c.blocks.size(); ++b) {
            const Block &p = c.blocks[a], &q = c.blocks[b];
            if (p.layer == q.layer && span(p.x0, p.x0 + p.w, q.x0, q.x0 + q.w) * span(p.y0, p.y0 + p.h, q.y0, q.y0 + q.h) > 0)
                return fail(TH_ERR_CASE, "power blocks '" + p.name + "' and '" + q.name +
                                             "' overlap (3D-ICE needs disjoint blocks)");
        }

    const double Tk = c.T_amb + 273.15;
    std::string stk;
    std::vector<std::pair<std::string, std::string>> extra;  // layout and floorplan files
    put(stk, "// Generated by thermal3d export3dice. Units: um, W/(um K), J/(um^3 K), K.\n\n");
    for (size_t L = 0; L < nl; ++L)
        put(stk, "material M%zu :   // %s\n   thermal conductivity     %.10e ;\n   volumetric heat capacity %.10e ;\n\n",
            L, c.layers[L].name.c_str(), c.layers[L].k * 1e-6, c.layers[L].rho_c * 1e-18);
    put(stk, "top heat sink :\n   heat transfer coefficient %.10e ;\n   temperature %.4f ;\n\n", c.h_top * 1e-12, Tk);
    put(stk, "dimensions :\n   chip length %.6f, width %.6f ;\n   cell length %.6f, width %.6f ;\n   non-uniform false ;\n\n",
        Lx, Ly, cell_um, cell_um);
    // Layer definitions (one per plane of an inactive layer), then the dies;
    // the stack section lists both top to bottom.
    std::vector<std::string> stack;  // bottom to top
    for (size_t L = 0; L < nl; ++L) {
        const Layer& ly = c.layers[L];
        if (ly.active) {
            std::string e;
            put(e, "   die   D%zu DIE%zu floorplan \"./L%zu.flp\" ;\n", L, L, L);
            stack.push_back(e);
            continue;
        }
        for (int p = 0; p < ly.planes; ++p) {
            put(stk, "layer LT%zu_%d :   // %s, plane %d\n   height %.6f ;\n   material M%zu ;\n", L, p, ly.name.c_str(),
                p, ly.thickness * 1e6 / ly.planes, L);
            const int l = first[L] + p;
            if (has_tsv(l)) {
                const std::string name = "LT" + std::to_string(L) + "_" + std::to_string(p) + ".lay";
                extra.emplace_back(name, layout_text(m, l, ly, cell_um));
                put(stk, "   layout \"./%s\" ;\n", name.c_str());
            }
            stk += "\n";
            std::string e;
            put(e, "   layer P%zu_%d LT%zu_%d ;\n", L, p, L, p);
            stack.push_back(e);
        }
    }
    for (size_t L = 0; L < nl; ++L)
        if (c.layers[L].active)
            put(stk, "die DIE%zu :   // %s\n   source %.6f M%zu ;\n\n", L, c.layers[L].name.c_str(),
                c.layers[L].thickness * 1e6, L);
    stk += "stack:\n";
    for (auto it = stack.rbegin(); it != stack.rend(); ++it) stk += *it;
    put(stk, "\nsolver:\n   steady ;\n   initial temperature %.4f ;\n   numofcores 1 ;\n\noutput:\n", Tk);
    for (size_t L = 0; L < nl; ++L) {
        if (!c.layers[L].active) continue;
        put(stk, "   Tmap ( D%zu, \"tmap_L%zu.txt\", final ) ;\n", L, L);
        put(stk, "   Tflp ( D%zu, \"tflp_max_L%zu.txt\", maximum, final ) ;\n", L, L);
        put(stk, "   Tflp ( D%zu, \"tflp_avg_L%zu.txt\", average, final ) ;\n", L, L);
    }
    // One floorplan per die: its power blocks, or one empty element (3D-ICE
    // needs at least one per die).
    std::vector<std::pair<std::string, std::string>> flps;
    for (size_t L = 0; L < nl; ++L) {
        if (!c.layers[L].active) continue;
        std::string flp;
        for (const Block& b : c.blocks)
            if ((size_t)b.layer == L)
                put(flp, "%s :\n   position %.6f, %.6f ;\n   dimension %.6f, %.6f ;\n   power values %.10e ;\n\n",
                    ident(b.name).c_str(), b.x0 * 1e6, b.y0 * 1e6, b.w * 1e6, b.h * 1e6, b.watts);
        if (flp.empty()) put(flp, "none :\n   position 0, 0 ;\n   dimension %.6f, %.6f ;\n   power values 0 ;\n\n", Lx, Ly);
        flps.emplace_back("L" + std::to_string(L) + ".flp", flp);
    }
    if (!write_file(outdir + "/case.stk", stk))
        return fail(TH_ERR_ARG, "cannot write " + outdir + "/case.stk (does the folder exist?)");
    for (const auto& f : extra)
        if (!write_file(outdir + "/" + f.first, f.second)) return fail(TH_ERR_ARG, "cannot write " + outdir + "/" + f.first);
    for (const auto& f : flps)
        if (!write_file(outdir + "/" + f.first, f.second)) return fail(TH_ERR_ARG, "cannot write floorplan file");
    if (nx_out) *nx_out = nx;
    if (ny_out) *ny_out = ny;
    return TH_OK;
}

// A Tmap file: rows of numbers, '%' lines are comments (read_3dice_tmap,
// case3d.h:387-400). All rows must have the same length.
int read_tmap(const std::string& path, std::vector<double>& vals, int& rows, int& cols) {
    std::ifstream f(path);
    if (!f) return fail(TH_ERR_ARG, "cannot open " + path);
    vals.clear();
    rows = cols = 0;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream ss(line);
        std::vector<double> r;
        double v;
        while (ss >> v) r.push_back(v);
        if (r.empty()) continue;
        if (rows == 0) cols = (int)r.size();
        else if ((int)r.size() != cols) return fail(TH_ERR_CASE, path + ": rows of different lengths");
        vals.insert(vals.end(), r.begin(), r.end());
        ++rows;
    }
    return TH_OK;
}

// A Tflp file: block names from the "% Time(s) name(K) ..." header, one
// value per block from the data line (read_3dice_tflp, main3d.cu:380-403).
int read_tflp(const std::string& path, std::vector<std::string>& names, std::vector<double>& vals) {
    std::ifstream f(path);
    if (!f) return fail(TH_ERR_ARG, "cannot open " + path);
    names.clear();
    vals.clear();
    std::string line;
    while (std::getline(f, line)) {
        if (line.compare(0, 6, "% Time") == 0) {
            std::istringstream ss(line.substr(1));
            std::string tok;
            ss >> tok;  // the time column
            while (ss >> tok) {
                const size_t k = tok.find("(K)");
                names.push_back(k == std::string::npos ? tok : tok.substr(0, k));
            }
        } else if (!line.empty() && line[0] != '%') {
            std::istringstream ss(line);
            double t, v;
            ss >> t;
            while (ss >> v) vals.push_back(v);
        }
    }
    if (names.empty() || names.size() != vals.size())
        return fail(TH_ERR_CASE, path + ": no block names, or not one value per block");
    return TH_OK;
}

}  // namespace th

using namespace th;

extern "C" {

int th_case_export_3dice(const char* dir, double cell_um, const char* outdir, int* nx, int* ny) {
    clear_error();
    if (!dir || !outdir) return fail(TH_ERR_ARG, "null argument");
    Case c;
    std::string err;
    if (!read_case(dir, c, err)) return fail(TH_ERR_CASE, err);
    return export_3dice(c, cell_um, outdir, nx, ny);
}

int th_3dice_read_tmap(const char* path, int* rows, int* cols, double* values, long capacity) {
    clear_error();
    if (!path || !rows || !cols) return fail(TH_ERR_ARG, "null argument");
    std::vector<double> v;
    int r = 0, c = 0;
    const int rc = read_tmap(path, v, r, c);
    if (rc) return rc;
    *rows = r;
    *cols = c;
    if (values) {
        if ((long)v.size() > capacity) return fail(TH_ERR_ARG, "tmap: values buffer too small");
        std::memcpy(values, v.data(), v.size() * sizeof(double));
    }
    return TH_OK;
}

int th_3dice_read_tflp(const char* path, char* names, int name_len, double* values, int capacity, int* count) {
    clear_error();
    if (!path || !count) return fail(TH_ERR_ARG, "null argument");
    std::vector<std::string> nm;
    std::vector<double> v;
    const int rc = read_tflp(path, nm, v);
    if (rc) return rc;
    *count = (int)v.size();
    if (!names && !values) return TH_OK;
    if ((int)v.size() > capacity) return fail(TH_ERR_ARG, "tflp: buffers too small");
    for (size_t k = 0; k < v.size(); ++k) {
        if (values) values[k] = v[k];
        if (names) {
            if ((int)nm[k].size() >= name_len) return fail(TH_ERR_ARG, "tflp: block name longer than name_len - 1");
            std::memcpy(names + k * (size_t)name_len, nm[k].c_str(), nm[k].size() + 1);
        }
    }
    return TH_OK;
}

}  // extern "C"
