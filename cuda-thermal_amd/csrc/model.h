// model.h - the layered stack model on the host: geometry, per-cell
// materials, face conductances, power and TSV maps (the role of the
// reference's src/stack3d.h and src/case3d.h), plus the error channel shared
// by the whole library.
#pragma once
#include <string>
#include <vector>

#include "../../include/thermal_hip.h"

namespace th {

// Thread-local last error; every C-ABI function reports through it.
int fail(int code, const std::string& msg);
void clear_error();

struct Layer {
    std::string name;
    double thickness;  // m
    double k;          // W/(m K)
    double rho_c;      // J/(m^3 K)
    int planes;        // z-planes in this layer
    bool active;       // may carry power
};

// The stack on an nx x ny lateral grid; per-cell arrays in the natural
// layout (x fastest, then y, then z), as the reference's Model3D
// (stack3d.h:45-79).
struct Model {
    int nx = 0, ny = 0, nz = 0;
    double Lx = 0, Ly = 0, dx = 0, dy = 0;
    double h_top = 0, h_bot = 0, T_amb = 25.0;
    double tsv_k = 400.0, tsv_rho_c = 3.45e6;  // copper fill
    std::vector<Layer> layers;
    std::vector<int> plane_layer;
    std::vector<double> dz;
    std::vector<double> tsv_frac, P;
    std::vector<float> gE, gN, gU, gB, cap;
    bool finalized = false;

    size_t nxy() const { return (size_t)nx * ny; }
    size_t ncells() const { return nxy() * nz; }
    size_t at(int i, int j, int l) const { return ((size_t)l * ny + j) * nx + i; }
};

// TSV effective medium with copper fraction f: vertical paths in parallel,
// lateral by the Rayleigh / Maxwell-Garnett rule for parallel cylinders
// (stack3d.h:20-25).
inline double kz_mix(double k, double kc, double f) { return f * kc + (1 - f) * k; }
inline double kxy_mix(double k, double kc, double f) {
    return k * ((1 + f) * kc + (1 - f) * k) / ((1 - f) * kc + (1 + f) * k);
}

int make_model(const std::vector<Layer>& layers, int nx, int ny, double Lx, double Ly,
               double h_top, double h_bot, Model& m);
std::vector<Layer> default_layers();
void finalize(Model& m);
void power_block(Model& m, int layer, double x0, double y0, double w, double h, double watts);
void tsv_region(Model& m, int layer, double x0, double y0, double w, double h, double f);
bool hotspot(Model& m, int layer, double xc, double yc, double size, double watts);
double total_power(const Model& m);
std::vector<double> chain_1d(const Model& m);
double heat_out(const Model& m, const double* T);
void block_temps(const Model& m, const double* T, int layer, double x0, double y0, double w, double h, double& tmax,
                 double& tavg);

// A case folder (case3d.h:1-13 format).
struct Block {
    std::string name;
    int layer;
    double x0, y0, w, h, watts;  // m, W
};
struct Tsv {
    int layer;
    double x0, y0, w, h, fraction;
};
struct Case {
    std::vector<Layer> layers;
    double Lx = 10e-3, Ly = 10e-3, h_top = 2e4, h_bot = 0.0, T_amb = 25.0;
    std::vector<Block> blocks;
    std::vector<Tsv> tsvs;
};
bool read_case(const std::string& dir, Case& c, std::string& err);
// 3D-ICE interchange (ice.cpp)
int export_3dice(const Case& c, double cell_um, const std::string& outdir, int* nx, int* ny);
int read_tmap(const std::string& path, std::vector<double>& vals, int& rows, int& cols);
int read_tflp(const std::string& path, std::vector<std::string>& names, std::vector<double>& vals);
Case builtin_example();
int case_model(const Case& c, int nx, int ny, bool with_tsv, Model& m);
// The surrogate's forward pass on a device input, queued on `stream` (a
// hipStream_t; surrogate.hip): *out receives its output field [H][W][C] in
// device memory, valid in that stream's order. The surrogate must live on
// `device`.
int surrogate_forward_on(th_surrogate* sur, const float* d_in, void* stream, int device, const float** out, int* H,
                         int* W, int* C);

}  // namespace th

struct th_model {
    th::Model m;
    std::vector<th::Block> blocks;  // power blocks of the case it was built from
    int tsv_regions = 0;
};
