// homo3d.hpp - C++ façade with the reference's host API names over the homo3d C ABI.
//
// Header-only; includes nothing from HIP.  Mirrors lavenklau/homo3d:
//   cfg::HomoConfig                       cmdline.h:35-60
//   homo::Homogenization                  homogenization/homogenization.h:17-61
//   homo::Grid / homo::MG (level views)   homogenization/grid.h:104-418, homogenization/MG.h:14-45
//   homo::OCOptimizer                     optimization/OCOptimizer.h:7-20
//   homo::MMAOptimizer                    optimization/mmaOptimizer.h:12-34
// Errors are thrown as homo::Homo3dError (a std::runtime_error, as the reference throws
// std::runtime_error("MG numerical explode") from MG::solveEquation, MG.cpp:153).
// Level operations act on all six macro-strain load cases at once.
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "homo3d.h"

namespace cfg {

enum class Objective : uint8_t { bulk, shear, npr, custom };
enum class Symmetry : uint8_t { reflect3, reflect6, rotate3, NONE };
enum class InitWay : uint8_t { random, randcenter, noise, manual, interp, rep_randcenter, P, G, D, IWP };

struct HomoConfig {
    Objective obj = Objective::bulk;
    Symmetry sym = Symmetry::reflect6;
    InitWay winit = InitWay::randcenter;
    int reso[3] = {64, 64, 64};
    double volRatio = 0.3;
    double youngsModulu = 10.0;
    double poissonRatio = 0.3;
    double finthres = 5e-4;
    double filterRadius = 2.0;
    double designStep = 0.05;
    double dampRatio = 0.5;
    double femRelThres = 1e-2;
    std::string outprefix;
    std::string testname = "none";
    std::string inputrho;
    bool useManagedMemory = true;
    int logrho = 0, logc = 0, logsens = 0, logobj = 0;
    int max_iter = 300;
    int initperiod = 10;
    // fields of this build (flags -solver, -seed, -periodfilt, -usesym, -optimizer)
    int solver = HOMO3D_SOLVER_MG;
    unsigned long long seed = 0;
    bool periodfilt = false;
    bool usesym = true;
    std::string optimizer = "oc";
    // gflags-compatible parser (cmdline.cpp:29-191): "-name=value", "--name=value" or "-name value".
    // Unknown flags and bad values throw std::invalid_argument.
    void parse(int argc, char** argv);
};

}  // namespace cfg

namespace homo {

class Homo3dError : public std::runtime_error {
public:
    Homo3dError(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline int sym_code(cfg::Symmetry s) {
    switch (s) {
        case cfg::Symmetry::reflect3: return HOMO3D_SYM_REFLECT3;
        case cfg::Symmetry::reflect6: return HOMO3D_SYM_REFLECT6;
        case cfg::Symmetry::rotate3: return HOMO3D_SYM_ROTATE3;
        default: return HOMO3D_SYM_NONE;
    }
}

inline homo3d_config to_c(const cfg::HomoConfig& c, int solver) {
    homo3d_config k;
    homo3d_default_config(&k);
    for (int i = 0; i < 3; i++) k.reso[i] = c.reso[i];
    k.youngs_modulus = c.youngsModulu;
    k.poisson_ratio = c.poissonRatio;
    k.fem_rel_thres = c.femRelThres;
    k.solver = solver;
    return k;
}

struct Homogenization;

// One level of the hierarchy (homo::Grid).  Level 0 is the matrix-free fine grid.
struct Grid {
    homo3d_handle* h = nullptr;
    int level = 0;
    int cellReso[3] = {0, 0, 0};
    bool is_root() const { return level == 0; }
    void chk(int rc, const char* what) const {
        if (rc) throw Homo3dError(rc, std::string(what) + ": " + homo3d_last_error(h));
    }
    void gs_relaxation(float w_SOR = 1.f, int times_ = 1) {
        for (int i = 0; i < times_; i++) chk(homo3d_smooth(h, level, w_SOR, 0), "gs_relaxation");
    }
    void update_residual() { chk(homo3d_residual(h, level), "update_residual"); }
    void restrict_residual() { chk(homo3d_restrict(h, level), "restrict_residual"); }
    void prolongate_correction() { chk(homo3d_prolong(h, level), "prolongate_correction"); }
    void enforce_unit_macro_strain() { chk(homo3d_macro_rhs(h), "enforce_unit_macro_strain"); }
    void solveHostEquation() { chk(homo3d_coarse_solve(h), "solveHostEquation"); }
    // v3_download / v3_upload of u, f or r (HOMO3D_FIELD_*), 6 x 3 x n^3 floats
    std::vector<float> download(int field) const {
        std::vector<float> v((size_t)18 * cellReso[0] * cellReso[1] * cellReso[2]);
        chk(homo3d_field_copy(h, level, field, v.data(), 0), "v3_download");
        return v;
    }
    void upload(int field, std::vector<float>& v) { chk(homo3d_field_copy(h, level, field, v.data(), 1), "v3_upload"); }
    void reset_displacement() {
        std::vector<float> z((size_t)18 * cellReso[0] * cellReso[1] * cellReso[2], 0.f);
        upload(HOMO3D_FIELD_U, z);
    }
};

// The hierarchy and its solvers (homo::MG).
struct MG {
    homo3d_handle* h = nullptr;
    std::vector<std::shared_ptr<Grid>> grids;
    std::shared_ptr<Grid> getRootGrid() { return grids.front(); }
    // MG::v_cycle(w_SOR, pre, post) (MG.cpp:46-70): forward GS sweeps with relaxation w_SOR.
    void v_cycle(float w_SOR = 1.f, int pre = 1, int post = 1) {
        int rc = homo3d_vcycle_sweeps(h, w_SOR, pre, post);
        if (rc) throw Homo3dError(rc, std::string("v_cycle: ") + homo3d_last_error(h));
    }
    // MG::solveEquation (MG.cpp:113-203): all six load cases; returns the worst relative residual.
    double solveEquation(double tol = 1e-2, bool with_guess = true) {
        homo3d_solve_stats st;
        int rc = homo3d_solve(h, tol, with_guess ? 1 : 0, &st);
        if (rc == HOMO3D_DIVERGED) throw Homo3dError(rc, "MG numerical explode");
        if (rc && rc != HOMO3D_NOT_CONVERGED) throw Homo3dError(rc, std::string("solveEquation: ") + homo3d_last_error(h));
        double w = 0;
        for (double r : st.relres) w = std::max(w, r);
        return w;
    }
    void updateStencils() {}  // stencils are rebuilt by Homogenization::update
};

struct Homogenization {
    explicit Homogenization(cfg::HomoConfig config_) : config(config_) {
        homo3d_config c = to_c(config, config.solver);
        int rc = homo3d_create(&c, &h_);
        if (rc) throw Homo3dError(rc, "Homogenization: unsupported configuration");
        int sizes[16];
        int nl = homo3d_levels(h_, sizes, 16);
        mg_ = std::make_unique<MG>();
        mg_->h = h_;
        for (int l = 0; l < nl; l++) {
            auto g = std::make_shared<Grid>();
            g->h = h_;
            g->level = l;
            g->cellReso[0] = g->cellReso[1] = g->cellReso[2] = sizes[l];
            mg_->grids.push_back(g);
        }
        grid = mg_->grids.front();
    }
    // Homogenization(int, int, int, double youngsModu = 1e6, double poissonRatio = 0.3) (homogenization.h:20)
    Homogenization(int xreso, int yreso, int zreso, double youngsModu = 1e6, double poissonRatio = 0.3)
        : Homogenization(make_cfg(xreso, yreso, zreso, youngsModu, poissonRatio)) {}
    ~Homogenization() {
        if (h_) homo3d_destroy(h_);
    }
    Homogenization(const Homogenization&) = delete;
    Homogenization& operator=(const Homogenization&) = delete;

    // Homogenization::update(float* rho, int pitchT): rho holds x = rho^p per cell (host, device or
    // managed memory); pitchT < 0 means unpitched.
    void update(float* rho = nullptr, int pitchT = -1) {
        if (!rho) return;
        chk(homo3d_update_density(h_, rho, pitchT < 0 ? grid->cellReso[0] : pitchT, -1), "update");
    }
    // Homogenization::elasticMatrix(double C[6][6]): solve the six load cases, then C^H.
    void elasticMatrix(double C[6][6]) {
        int rc = homo3d_solve(h_, config.femRelThres, 1, &last_stats);
        if (rc == HOMO3D_DIVERGED) throw Homo3dError(rc, "MG numerical explode");
        if (rc && rc != HOMO3D_NOT_CONVERGED) chk(rc, "elasticMatrix");
        double c[36];
        chk(homo3d_elastic_matrix(h_, c), "elasticMatrix");
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) C[i][j] = c[i * 6 + j];
    }
    void elasticMatrix(float C[6][6]) {
        double c[6][6];
        elasticMatrix(c);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) C[i][j] = (float)c[i][j];
    }
    // Homogenization::elasticMatrix(int i, int j) (homogenization.h:30): C^H_ij from the stored solutions.
    double elasticMatrix(int i, int j) {
        if (i < 0 || i > 5 || j < 0 || j > 5) throw Homo3dError(HOMO3D_BAD_ARG, "elasticMatrix: index out of range");
        double c[36];
        chk(homo3d_elastic_matrix(h_, c), "elasticMatrix");
        return c[i * 6 + j];
    }
    // Homogenization::elasticMatrix(float* rho, int i, int j) (homogenization.h:32): update with x = rho
    // (unpitched), solve the six load cases, return C^H_ij.
    double elasticMatrix(float* rho, int i, int j) {
        update(rho);
        double C[6][6];
        elasticMatrix(C);
        return C[i][j];
    }
    // Homogenization::Sensitivity(dC, sens, pitchT, lexiOrder) (homogenization.h:38).  lexiOrder = false
    // writes the library's own cell order (parity-blocked, n^3 unpitched; see homo3d_sensitivity_order).
    void Sensitivity(float dC[6][6], float* sens, int pitchT, bool lexiOrder = false) {
        chk(homo3d_sensitivity_order(h_, &dC[0][0], sens, pitchT < 0 ? grid->cellReso[0] : pitchT, lexiOrder ? 1 : 0,
                                     -1),
            "Sensitivity");
    }
    // Homogenization::Sensitivity(int i, int j, float* sens) (homogenization.h:34): dC^H_ij/dx per cell
    // from the stored solutions, lexicographic and unpitched.
    void Sensitivity(int i, int j, float* sens) {
        if (i < 0 || i > 5 || j < 0 || j > 5) throw Homo3dError(HOMO3D_BAD_ARG, "Sensitivity: index out of range");
        float dC[6][6] = {};
        dC[i][j] = 1.f;
        Sensitivity(dC, sens, grid->cellReso[0], true);
    }
    // Homogenization::Sensitivity(float* rho, int i, int j, float* sens) (homogenization.h:36): update
    // with x = rho, solve, then Sensitivity(i, j, sens).
    void Sensitivity(float* rho, int i, int j, float* sens) {
        double C[6][6];
        update(rho);
        elasticMatrix(C);
        Sensitivity(i, j, sens);
    }
    std::shared_ptr<Grid> getGrid() { return grid; }
    MG& mg() { return *mg_; }
    homo3d_handle* handle() { return h_; }

    cfg::HomoConfig config;
    std::shared_ptr<Grid> grid;
    std::unique_ptr<MG> mg_;
    homo3d_solve_stats last_stats{};

private:
    static cfg::HomoConfig make_cfg(int x, int y, int z, double E, double nu) {
        cfg::HomoConfig c;
        c.reso[0] = x;
        c.reso[1] = y;
        c.reso[2] = z;
        c.youngsModulu = E;
        c.poissonRatio = nu;
        return c;
    }
    void chk(int rc, const char* what) {
        if (rc) throw Homo3dError(rc, std::string(what) + ": " + homo3d_last_error(h_));
    }
    homo3d_handle* h_ = nullptr;
};

// homo::OCOptimizer: filterSens and update on the GPU (density-only handle per grid size).
struct OCOptimizer {
    int ne = 0;
    float step_limit = 0.02f;
    float damp = 0.5f;
    float minRho = 0.001f;
    bool periodic = false;
    OCOptimizer(int Ne, float min_density, float stepLimit, float dampExponent)
        : ne(Ne), step_limit(stepLimit), damp(dampExponent), minRho(min_density) {}
    OCOptimizer(float min_density, float stepLimit, float dampExponent)
        : step_limit(stepLimit), damp(dampExponent), minRho(min_density) {}
    ~OCOptimizer() {
        if (h_) homo3d_destroy(h_);
    }
    OCOptimizer(const OCOptimizer&) = delete;
    OCOptimizer& operator=(const OCOptimizer&) = delete;

    // filterSens(float* sens, const float* rho, size_t pitchT, int reso[3], float radius)
    void filterSens(float* sens, const float* rho, size_t pitchT, int reso[3], float radius = 2) {
        bind(reso[0]);
        chk(homo3d_oc_filter(h_, sens, rho, (int)pitchT, radius, periodic ? 1 : 0, -1), "filterSens");
    }
    // update(const float* sens, float* rho, float volratio) on an n^3 grid (n from Ne or a filterSens call)
    void update(const float* sens, float* rho, float volratio) {
        if (!h_) bind((int)std::lround(std::cbrt((double)ne)));
        chk(homo3d_oc_update(h_, sens, rho, n_, volratio, minRho, step_limit, damp, -1), "update");
    }

private:
    void bind(int n) {
        if (h_ && n == n_) return;
        if (h_) homo3d_destroy(h_);
        homo3d_config c;
        homo3d_default_config(&c);
        c.reso[0] = c.reso[1] = c.reso[2] = n;
        c.solver = HOMO3D_SOLVER_NONE;
        int rc = homo3d_create(&c, &h_);
        if (rc) throw Homo3dError(rc, "OCOptimizer: cannot create a density handle");
        n_ = n;
    }
    void chk(int rc, const char* what) {
        if (rc) throw Homo3dError(rc, std::string("OCOptimizer::") + what + ": " + homo3d_last_error(h_));
    }
    homo3d_handle* h_ = nullptr;
    int n_ = 0;
};

// homo::MMAOptimizer(int m, int n, double a0, double a, double c, double d)
struct MMAOptimizer {
    int n, m;
    double a0, a, c, d;
    MMAOptimizer(int nconstrain, int nvar, double a0_, double a_, double c_, double d_)
        : n(nvar), m(nconstrain), a0(a0_), a(a_), c(c_), d(d_) {
        int rc = homo3d_mma_create(m, n, a0, a, c, d, 0, &h_);
        if (rc) throw Homo3dError(rc, "MMAOptimizer: create failed");
    }
    ~MMAOptimizer() {
        if (h_) homo3d_mma_destroy(h_);
    }
    MMAOptimizer(const MMAOptimizer&) = delete;
    MMAOptimizer& operator=(const MMAOptimizer&) = delete;
    void setBound(float xmin, float xmax) { chk(homo3d_mma_set_bound(h_, xmin, xmax), "setBound"); }
    // setBound(float* xmin, float* xmax) (mmaOptimizer.h:30): n bounds each, host or device memory
    void setBound(float* xmin, float* xmax) { chk(homo3d_mma_set_bounds(h_, xmin, xmax, -1), "setBound"); }
    // update(int itn, float* x, float* dfdx, float* gval, float** dgdx) (mmaOptimizer.h:34)
    void update(int itn, float* x, float* dfdx, float* gval, float** dgdx) {
        chk(homo3d_mma_update(h_, itn, x, dfdx, gval, (const float* const*)dgdx, -1), "update");
    }

private:
    void chk(int rc, const char* what) {
        if (rc) throw Homo3dError(rc, std::string("MMAOptimizer::") + what + ": " + homo3d_mma_last_error(h_));
    }
    homo3d_mma* h_ = nullptr;
};

// Optimisation drivers (Framework.cu:137-366, homogenization/testAutoDiff.cu:571-607), implemented in
// libhomo3d_framework.so.  Each returns the objective history; outputs go to config.outprefix.
std::vector<double> example_opti_bulk(cfg::HomoConfig config);
std::vector<double> example_opti_npr(cfg::HomoConfig config);
std::vector<double> example_opti_shear_isotropy(cfg::HomoConfig config);
std::vector<double> example_opti_npr_mma(cfg::HomoConfig config);
void example_yours(cfg::HomoConfig config);  // weak user hook, replaced by a strong definition
void runCustom(cfg::HomoConfig config);
std::vector<double> runInstance(cfg::HomoConfig config);

}  // namespace homo

// The reference's main (main.cpp:22-48) as a C entry point: 0 on success, 1 on a flag or I/O error,
// 2 on a solver error.
extern "C" int homo3d_framework_main(int argc, char** argv);
