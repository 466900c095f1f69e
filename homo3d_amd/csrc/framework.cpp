// Optimisation drivers on the C++ façade (Framework.cu:61-366, homogenization/Framework.cuh:221-343,
// homogenization/testAutoDiff.cu:282-607, cmdline.cpp:4-191).  Pure host C++: every GPU operation goes
// through the homo3d C ABI via include/homo3d.hpp.
//
// The expression layer of the reference (TensorVar, Ch(i,j) autodiff) is replaced by explicit
// objectives: each driver forms dJ/dC^H by hand, calls Homogenization::Sensitivity and applies the
// chain rule of x = rho^3.  The density fields of the loops (rho, x, sensitivities, gradients) stay in
// device memory, as the reference's tensors do; only C^H, the objective and constraint values and the
// logged densities cross to the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/homo3d.hpp"

namespace {

std::string out_path(const cfg::HomoConfig& c, const std::string& name) { return c.outprefix + name; }

// gflags-style value parsing
bool parse_bool(const std::string& v) {
    if (v == "" || v == "1" || v == "true" || v == "True" || v == "yes") return true;
    if (v == "0" || v == "false" || v == "False" || v == "no") return false;
    throw std::invalid_argument("bad bool flag value " + v);
}

}  // namespace

// cfg::HomoConfig::parse (cmdline.cpp:29-191) with the reference's flag names and defaults
// (cmdline.cpp:4-27), without gflags: -name=value, --name=value or -name value.
void cfg::HomoConfig::parse(int argc, char** argv) {
    std::map<std::string, std::string> kv;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a.size() < 2 || a[0] != '-') throw std::invalid_argument("unexpected argument " + a);
        a = a.substr(a[1] == '-' ? 2 : 1);
        auto eq = a.find('=');
        if (eq != std::string::npos) {
            kv[a.substr(0, eq)] = a.substr(eq + 1);
        } else if (i + 1 < argc && argv[i + 1][0] != '-') {
            kv[a] = argv[++i];
        } else {
            kv[a] = "";
        }
    }
    std::ostringstream dump;
    for (auto& [k, v] : kv) {
        dump << "--" << k << "=" << v << "\n";
        if (k == "reso") reso[0] = reso[1] = reso[2] = std::stoi(v);
        else if (k == "obj") {
            if (v == "bulk") obj = Objective::bulk;
            else if (v == "shear") obj = Objective::shear;
            else if (v == "npr") obj = Objective::npr;
            else if (v == "custom") obj = Objective::custom;
            else throw std::invalid_argument("unrecognized Objective type " + v);
        } else if (k == "init") {
            if (v == "rand") winit = InitWay::random;
            else if (v == "randc") winit = InitWay::randcenter;
            else if (v == "reprandc") winit = InitWay::rep_randcenter;
            else if (v == "noise") winit = InitWay::noise;
            else if (v == "interp") winit = InitWay::interp;
            else if (v == "P") winit = InitWay::P;
            else if (v == "G") winit = InitWay::G;
            else if (v == "D") winit = InitWay::D;
            else if (v == "IWP") winit = InitWay::IWP;
            else if (v == "manual") winit = InitWay::manual;
            else throw std::invalid_argument("unrecognized density initialize mean " + v);
        } else if (k == "sym") {
            if (v == "reflect3") sym = Symmetry::reflect3;
            else if (v == "reflect6") sym = Symmetry::reflect6;
            else if (v == "rotate3") sym = Symmetry::rotate3;
            else if (v == "none") sym = Symmetry::NONE;
            else throw std::invalid_argument("unrecognized symmetry type " + v);
        } else if (k == "vol") volRatio = std::stod(v);
        else if (k == "E") youngsModulu = std::stod(v);
        else if (k == "mu") poissonRatio = std::stod(v);
        else if (k == "prefix") outprefix = v;
        else if (k == "logrho") logrho = std::stoi(v);
        else if (k == "logc") logc = std::stoi(v);
        else if (k == "logsens") logsens = std::stoi(v);
        else if (k == "logobj") logobj = std::stoi(v);
        else if (k == "test") testname = v;
        else if (k == "managedmem") useManagedMemory = parse_bool(v);
        else if (k == "in") inputrho = v;
        else if (k == "N") max_iter = std::stoi(v);
        else if (k == "initperiod") initperiod = std::stoi(v);
        else if (k == "finthres") finthres = std::stod(v);
        else if (k == "filter") filterRadius = std::stod(v);
        else if (k == "step") designStep = std::stod(v);
        else if (k == "damp") dampRatio = std::stod(v);
        else if (k == "relthres") femRelThres = std::stod(v);
        else if (k == "periodfilt") periodfilt = parse_bool(v);
        else if (k == "usesym") usesym = parse_bool(v);
        else if (k == "solver") {
            if (v == "mg") solver = HOMO3D_SOLVER_MG;
            else if (v == "pcg") solver = HOMO3D_SOLVER_PCG;
            else throw std::invalid_argument("unrecognized solver " + v);
        } else if (k == "seed") seed = std::stoull(v);
        else if (k == "optimizer") {
            if (v != "oc" && v != "mma") throw std::invalid_argument("unrecognized optimizer " + v);
            optimizer = v;
        } else throw std::invalid_argument("unknown flag -" + k);
    }
    std::ofstream ofs(outprefix + "cmdline");
    ofs << dump.str();
}

namespace homo {

// symmetrizeField (Framework.cuh:221-235) on a density-only handle
struct DensityOps {
    homo3d_handle* h = nullptr;
    int n = 0;
    explicit DensityOps(int n_) : n(n_) {
        homo3d_config c;
        homo3d_default_config(&c);
        c.reso[0] = c.reso[1] = c.reso[2] = n;
        c.solver = HOMO3D_SOLVER_NONE;
        int rc = homo3d_create(&c, &h);
        if (rc) throw Homo3dError(rc, "density handle");
    }
    ~DensityOps() {
        if (h) homo3d_destroy(h);
    }
    void chk(int rc, const char* what) const {
        if (rc) throw Homo3dError(rc, std::string(what) + ": " + homo3d_last_error(h));
    }
    // n^3 arrays with pitch n, in host or device memory (the C ABI tells them apart)
    void symmetrize(float* f, cfg::Symmetry s, bool usesym) {
        if (!usesym || s == cfg::Symmetry::NONE) return;
        chk(homo3d_symmetrize(h, f, n, sym_code(s), -1), "symmetrize");
    }
    void symmetrize(std::vector<float>& f, cfg::Symmetry s, bool usesym) { symmetrize(f.data(), s, usesym); }
    // conv(radial_convker_t<Spline4>(radius)) (non-periodic, normalised) or its adjoint
    void spline4_filter(const float* in, float* out, double radius, bool backward) {
        chk(homo3d_density_filter(h, in, out, n, radius, 0, 0, backward ? 1 : 0, -1), "density filter");
    }
    // sgm(k = 30, c = 0.5) or its backward (y the forward output, in = dJ/dy)
    void sgm(const float* in, const float* y, float* out) {
        chk(homo3d_projection(h, in, y, out, n, 30.0, 0.5, y ? 1 : 0, -1), "projection");
    }
    // x = rho^p and drho = p rho^(p-1) dx
    void pow(const float* rho, float* x, double p) { chk(homo3d_density_pow(h, rho, x, n, p, -1), "density pow"); }
    void pow_grad(const float* rho, const float* dx, float* drho, double p) {
        chk(homo3d_density_pow_grad(h, rho, dx, drho, n, p, -1), "density pow grad");
    }
    double sum(const float* a) {
        double t = 0;
        chk(homo3d_field_sum(h, a, n, &t, -1), "field sum");
        return t;
    }
    void fill(float* a, float v) { chk(homo3d_field_fill(h, a, n, v, -1), "field fill"); }
};

// An n^3 float array in device memory (the C ABI's allocation calls: this file needs no HIP headers).
struct DevField {
    float* p = nullptr;
    size_t count = 0;
    explicit DevField(size_t c = 0) { resize(c); }
    DevField(const DevField&) = delete;
    DevField& operator=(const DevField&) = delete;
    DevField(DevField&& o) noexcept : p(o.p), count(o.count) {
        o.p = nullptr;
        o.count = 0;
    }
    ~DevField() {
        if (p) homo3d_device_free(p);
    }
    void resize(size_t c) {
        if (c == count) return;
        if (p) homo3d_device_free(p);
        p = nullptr;
        count = c;
        if (!c) return;
        void* q = nullptr;
        int rc = homo3d_device_alloc(&q, sizeof(float) * c);
        if (rc) throw Homo3dError(rc, "device allocation of " + std::to_string(c) + " floats");
        p = static_cast<float*>(q);
    }
    void upload(const std::vector<float>& h) {
        resize(h.size());
        int rc = homo3d_device_copy(p, h.data(), sizeof(float) * count, 0);
        if (rc) throw Homo3dError(rc, "upload");
    }
    std::vector<float> download() const {
        std::vector<float> h(count);
        int rc = homo3d_device_copy(h.data(), p, sizeof(float) * count, 1);
        if (rc) throw Homo3dError(rc, "download");
        return h;
    }
};

// ConvergeChecker (Framework.cuh:312-343): converged when the relative change stays below the
// threshold for three consecutive iterations, checked after iteration 11.
struct ConvergeChecker {
    static constexpr int kLen = 10;
    double thres;
    double obj[kLen] = {};
    int p = 0;
    explicit ConvergeChecker(double t = 1e-3) : thres(t) {}
    double get(int id) const {
        while (id < 0) id += kLen;
        return obj[id % kLen];
    }
    bool is_converge(int iter, double val) {
        p = iter % kLen;
        obj[p] = val;
        if (iter <= 11) return false;
        for (int i = 0; i < 3; i++)
            if (std::fabs(get(p - i) - get(p - i - 1)) / std::fabs(get(p - i - 1)) > thres) return false;
        return true;
    }
};

static float tanproj(float v, float beta, float tau = 0.5f) {
    const float t = std::tanh(beta * tau);
    return (t + std::tanh(beta * (v - tau))) / (t + std::tanh(beta * (1.f - tau)));
}

// initDensity (Framework.cu:61-134).  randc / rand / reprandc use the seeded generator;
// manual reads n^3 raw float32 values from config.inputrho (OpenVDB is not available).
void initDensity(std::vector<float>& rho, const cfg::HomoConfig& c, DensityOps& ops) {
    const int n = c.reso[0];
    const size_t N = (size_t)n * n * n;
    rho.assign(N, 0.f);
    const float pi = 3.1415926f;
    auto field = [&](auto fn) {
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    const float p[3] = {float(i) / n, float(j) / n, float(k) / n};
                    rho[i + (size_t)n * (j + (size_t)n * k)] = std::max(std::min(fn(p), 1.f), 0.001f);
                }
    };
    switch (c.winit) {
        case cfg::InitWay::random:
        case cfg::InitWay::randcenter:
        case cfg::InitWay::rep_randcenter: {
            int rc = homo3d_randc_density(n, c.seed, c.volRatio, c.usesym ? sym_code(c.sym) : HOMO3D_SYM_NONE,
                                          rho.data());
            if (rc) throw Homo3dError(rc, "randc density");
            break;
        }
        case cfg::InitWay::manual:
        case cfg::InitWay::interp: {
            std::ifstream ifs(c.inputrho, std::ios::binary);
            if (!ifs.read(reinterpret_cast<char*>(rho.data()), sizeof(float) * N))
                throw Homo3dError(HOMO3D_BAD_ARG, "cannot read " + std::to_string(N) + " floats from " + c.inputrho);
            break;
        }
        case cfg::InitWay::noise: {
            std::mt19937_64 gen(c.seed);
            std::uniform_real_distribution<float> U(0.f, 1.f);
            for (auto& v : rho) v = U(gen);
            ops.symmetrize(rho, c.sym, c.usesym);
            for (auto& v : rho) v = tanproj(v, 20.f, 0.5f);
            double s = 0;
            for (float v : rho) s += v;
            const float sc = (float)(c.volRatio / (s / N));
            for (auto& v : rho) v = std::min(std::max(v * sc, 0.001f), 1.f);
            break;
        }
        case cfg::InitWay::P:
            field([&](const float* p) {
                return tanproj(-(std::cos(2 * pi * p[0]) + std::cos(2 * pi * p[1]) + std::cos(2 * pi * p[2])), 20);
            });
            break;
        case cfg::InitWay::G:
            field([&](const float* p) {
                float s[3], co[3];
                for (int i = 0; i < 3; i++) {
                    s[i] = std::sin(2 * pi * p[i]);
                    co[i] = std::cos(2 * pi * p[i]);
                }
                return tanproj(s[0] * co[1] + s[2] * co[0] + s[1] * co[2], 20);
            });
            break;
        case cfg::InitWay::D:
            field([&](const float* p) {
                const float x = p[0], y = p[1], z = p[2];
                return tanproj(std::cos(2 * pi * x) * std::cos(2 * pi * y) * std::cos(2 * pi * z) -
                                   std::sin(2 * pi * x) * std::sin(2 * pi * y) * std::sin(2 * pi * z),
                               20);
            });
            break;
        case cfg::InitWay::IWP:
            field([&](const float* p) {
                const float x = p[0], y = p[1], z = p[2];
                return tanproj(2 * (std::cos(2 * pi * x) * std::cos(2 * pi * y) + std::cos(2 * pi * y) * std::cos(2 * pi * z) +
                                    std::cos(2 * pi * z) * std::cos(2 * pi * x)) -
                                   (std::cos(4 * pi * x) + std::cos(4 * pi * y) + std::cos(4 * pi * z)),
                               20);
            });
            break;
    }
    ops.symmetrize(rho, c.sym, c.usesym);
    for (auto& v : rho) v = std::min(std::max(v, 0.001f), 1.f);
}

// logIter (Framework.cu:7-58): C every 5 iterations (36 float32, row-major), Clog/objlog text logs,
// rho as raw float32 (rho_XXXX when logrho divides the iteration).
void logIter(int iter, const cfg::HomoConfig& c, const DevField& rho_dev, const double C[6][6], double obj) {
    std::vector<float> rho;  // read back only when a density file is due
    auto write_C = [&](const std::string& f) {
        float Cf[36];
        for (int i = 0; i < 36; i++) Cf[i] = (float)C[i / 6][i % 6];
        std::ofstream(f, std::ios::binary).write(reinterpret_cast<const char*>(Cf), sizeof(Cf));
    };
    auto write_rho = [&](const std::string& f) {
        if (rho.empty()) rho = rho_dev.download();
        std::ofstream(f, std::ios::binary).write(reinterpret_cast<const char*>(rho.data()), sizeof(float) * rho.size());
    };
    if (iter % 5 == 0) {
        write_rho(out_path(c, "rho"));
        write_C(out_path(c, "C"));
    }
    if (c.logrho && iter % c.logrho == 0) {
        char b[64];
        std::snprintf(b, sizeof(b), "rho_%04d", iter);
        write_rho(out_path(c, b));
    }
    if (c.logc && iter % c.logc == 0) {
        std::ofstream ofs(out_path(c, "Clog"), iter == 0 ? std::ios::trunc : std::ios::app);
        ofs << "iter " << iter << " ";
        for (int i = 0; i < 36; i++) ofs << C[i / 6][i % 6] << " ";
        ofs << "\n";
    }
    if (c.logobj && iter % c.logobj == 0) {
        std::ofstream ofs(out_path(c, "objlog"), iter == 0 ? std::ios::trunc : std::ios::app);
        ofs << "iter " << iter << " obj = " << obj << "\n";
    }
}

// One forward/backward pass: x = rho^3 -> C^H, then dJ/drho = 3 rho^2 * dJ/dx for dJ/dC = dC.
struct Evaluator {
    Homogenization& hom;
    DensityOps& ops;
    DevField x, sens;
    Evaluator(Homogenization& h, DensityOps& o) : hom(h), ops(o) {}
    void forward(const DevField& rho, double C[6][6]) {
        x.resize(rho.count);
        ops.pow(rho.p, x.p, 3.0);
        hom.update(x.p, hom.grid->cellReso[0]);
        hom.elasticMatrix(C);
    }
    void backward(const DevField& rho, float dC[6][6], DevField& grad) {
        sens.resize(rho.count);
        hom.Sensitivity(dC, sens.p, hom.grid->cellReso[0], true);
        grad.resize(rho.count);
        ops.pow_grad(rho.p, sens.p, grad.p, 3.0);
    }
};

using ObjectiveFn = std::function<double(int iter, const double C[6][6], float dC[6][6])>;

// OC loop shared by example_opti_bulk and example_opti_npr (Framework.cu:137-279).
std::vector<double> run_oc(cfg::HomoConfig config, const ObjectiveFn& objective) {
    Homogenization hom(config);
    DensityOps ops(config.reso[0]);
    DevField rho, grad;
    {
        std::vector<float> r0;
        initDensity(r0, config, ops);
        std::ofstream(out_path(config, "initRho"), std::ios::binary)
            .write(reinterpret_cast<const char*>(r0.data()), sizeof(float) * r0.size());
        rho.upload(r0);
    }
    Evaluator ev(hom, ops);
    OCOptimizer oc(0.001f, (float)config.designStep, (float)config.dampRatio);
    oc.periodic = config.periodfilt;
    ConvergeChecker criteria(config.finthres);
    std::vector<double> objlist;
    int reso[3] = {config.reso[0], config.reso[1], config.reso[2]};
    double C[6][6];
    for (int iter = 0; iter < config.max_iter; iter++) {
        ev.forward(rho, C);
        float dC[6][6] = {};
        const double val = objective(iter, C, dC);
        objlist.push_back(val);
        ev.backward(rho, dC, grad);
        std::printf(" * Iter %d   obj = %.4e   (solve %.3fs, iters %d..%d)\n", iter, val, hom.last_stats.seconds,
                    *std::min_element(hom.last_stats.iters, hom.last_stats.iters + 6),
                    *std::max_element(hom.last_stats.iters, hom.last_stats.iters + 6));
        if (criteria.is_converge(iter, val)) {
            std::printf("= converged\n");
            break;
        }
        ops.symmetrize(grad.p, config.sym, config.usesym);
        oc.filterSens(grad.p, rho.p, reso[0], reso, (float)config.filterRadius);
        oc.update(grad.p, rho.p, (float)config.volRatio);
        ops.symmetrize(rho.p, config.sym, config.usesym);
        logIter(iter, config, rho, C, val);
    }
    {
        const std::vector<float> rh = rho.download();
        std::ofstream(out_path(config, "rho"), std::ios::binary)
            .write(reinterpret_cast<const char*>(rh.data()), sizeof(float) * rh.size());
    }
    float Cf[36];
    for (int i = 0; i < 36; i++) Cf[i] = (float)C[i / 6][i % 6];
    std::ofstream(out_path(config, "C"), std::ios::binary).write(reinterpret_cast<const char*>(Cf), sizeof(Cf));
    std::ofstream ol(out_path(config, "objlist"));
    for (double v : objlist) ol << v << "\n";
    return objlist;
}

// example_opti_bulk (Framework.cu:137-200): maximise the bulk modulus with OC.
std::vector<double> example_opti_bulk(cfg::HomoConfig config) {
    return run_oc(config, [](int, const double C[6][6], float dC[6][6]) {
        const double v = -(C[0][0] + C[1][1] + C[2][2] + (C[0][1] + C[0][2] + C[1][2]) * 2) / 9.0;
        dC[0][0] = dC[1][1] = dC[2][2] = -1.f / 9.f;
        dC[0][1] = dC[0][2] = dC[1][2] = -2.f / 9.f;
        return v;
    });
}

// example_opti_npr (Framework.cu:202-279): negative Poisson ratio with a continuation factor 0.8^iter.
std::vector<double> example_opti_npr(cfg::HomoConfig config) {
    return run_oc(config, [](int iter, const double C[6][6], float dC[6][6]) {
        const double beta = std::pow(0.8, iter);
        const double v = C[0][1] + C[0][2] + C[1][2] - (C[0][0] + C[1][1] + C[2][2]) * beta;
        dC[0][1] = dC[0][2] = dC[1][2] = 1.f;
        dC[0][0] = dC[1][1] = dC[2][2] = (float)-beta;
        return v;
    });
}

// MMA loop with a volume constraint (the MMA usage of example_opti_shear_isotropy,
// Framework.cu:282-356; MMAOptimizer(2, ne, 1, 0, 1000, 1), setBound(0.001, 1)).  constraints(C, dC)
// returns extra constraint values and fills their dJ/dC (at most 3).
struct MmaConstraint {
    double value;
    float dC[6][6];
};
using ConstraintFn = std::function<std::vector<MmaConstraint>(const double C[6][6])>;

std::vector<double> run_mma(cfg::HomoConfig config, const ObjectiveFn& objective, const ConstraintFn& extra) {
    Homogenization hom(config);
    DensityOps ops(config.reso[0]);
    DevField rho, grad;
    {
        std::vector<float> r0;
        initDensity(r0, config, ops);
        std::ofstream(out_path(config, "initRho"), std::ios::binary)
            .write(reinterpret_cast<const char*>(r0.data()), sizeof(float) * r0.size());
        rho.upload(r0);
    }
    Evaluator ev(hom, ops);
    const int ne = (int)rho.count;
    double C[6][6];
    ev.forward(rho, C);
    const int m = 1 + (int)extra(C).size();
    MMAOptimizer mma(m, ne, 1, 0, 1000, 1);
    mma.setBound(0.001f, 1.f);
    ConvergeChecker criteria(config.finthres);
    std::vector<double> objlist;
    std::vector<DevField> dg;
    for (int i = 0; i < m; i++) dg.emplace_back(ne);
    ops.fill(dg[0].p, (float)(1000.0 / ne));  // the volume constraint's gradient
    for (int iter = 0; iter < config.max_iter; iter++) {
        if (iter > 0) ev.forward(rho, C);
        float dC[6][6] = {};
        const double val = objective(iter, C, dC);
        objlist.push_back(val);
        ev.backward(rho, dC, grad);
        std::printf(" * Iter %d   obj = %.4e   (solve %.3fs)\n", iter, val, hom.last_stats.seconds);
        if (criteria.is_converge(iter, val)) {
            std::printf("= converged\n");
            break;
        }
        ops.symmetrize(grad.p, config.sym, config.usesym);
        // volume constraint g0 = (mean(rho) - vol) * 1000
        const double vsum = ops.sum(rho.p);
        std::vector<float> gval(m);
        gval[0] = (float)((vsum / ne - config.volRatio) * 1000.0);
        auto ex = extra(C);
        for (size_t k = 0; k < ex.size(); k++) {
            gval[1 + k] = (float)ex[k].value;
            ev.backward(rho, ex[k].dC, dg[1 + k]);
            ops.symmetrize(dg[1 + k].p, config.sym, config.usesym);
        }
        std::vector<float*> dgp(m);
        for (int i = 0; i < m; i++) dgp[i] = dg[i].p;
        mma.update(iter, rho.p, grad.p, gval.data(), dgp.data());
        ops.symmetrize(rho.p, config.sym, config.usesym);
        logIter(iter, config, rho, C, val);
    }
    {
        const std::vector<float> rh = rho.download();
        std::ofstream(out_path(config, "rho"), std::ios::binary)
            .write(reinterpret_cast<const char*>(rh.data()), sizeof(float) * rh.size());
    }
    std::ofstream ol(out_path(config, "objlist"));
    for (double v : objlist) ol << v << "\n";
    return objlist;
}

// example_opti_shear_isotropy (Framework.cu:282-356): maximise the shear modulus with an isotropy
// (Zener ratio) constraint and a volume constraint, by MMA.  The reference's Spline4 density filter
// and sigmoid in front of rho^3 are not applied here.
// Objective and isotropy constraint of example_opti_shear_isotropy (Framework.cu:300-323) for a given C.
// The constraint value is scale * (scale (A/B - 1)^2) - 0.1 with A = 2 (C33 + C44 + C55) and
// B = C00 + C11 + C22 - C01 - C02 - C12, while its gradient is that of the expression the reference
// back-propagates, scale (A/B - 1)^2: one factor of scale less, as there.
double shear_objective(const double C[6][6], float dC[6][6]) {
    dC[3][3] = dC[4][4] = dC[5][5] = -1.f / 3.f;
    return -(C[3][3] + C[4][4] + C[5][5]) / 3.0;
}

MmaConstraint shear_isotropy_constraint(const double C[6][6]) {
    const double A = (C[3][3] + C[4][4] + C[5][5]) * 2, B = C[0][0] + C[1][1] + C[2][2] - C[0][1] - C[0][2] - C[1][2];
    const double r = A / B - 1.0, scale = 1000.0;
    MmaConstraint k{};
    k.value = r * r * scale * scale - 0.1;
    const double dr = 2 * r * scale;
    const double dA = dr / B, dB = -dr * A / (B * B);
    k.dC[3][3] = k.dC[4][4] = k.dC[5][5] = (float)(2 * dA);
    k.dC[0][0] = k.dC[1][1] = k.dC[2][2] = (float)dB;
    k.dC[0][1] = k.dC[0][2] = k.dC[1][2] = (float)-dB;
    return k;
}

// example_opti_shear_isotropy (Framework.cu:282-356): maximise the shear modulus by MMA with a volume
// constraint and an isotropy (Zener ratio) constraint on the density rhop = sgm(conv(rho, Spline4(r)))^3.
// As there: the volume is that of sgm(conv(rho)); the constraint values carry a factor 1000 their
// gradients do not (the reference back-propagates the unscaled expressions); the objective and isotropy
// gradients are symmetrised, the volume gradient is not.
std::vector<double> example_opti_shear_isotropy(cfg::HomoConfig config) {
    Homogenization hom(config);
    DensityOps ops(config.reso[0]);
    DevField rho;
    {
        std::vector<float> r0;
        initDensity(r0, config, ops);
        std::ofstream(out_path(config, "initRho"), std::ios::binary)
            .write(reinterpret_cast<const char*>(r0.data()), sizeof(float) * r0.size());
        rho.upload(r0);
    }
    const int ne = (int)rho.count;
    MMAOptimizer mma(2, ne, 1, 0, 1000, 1);
    mma.setBound(0.001f, 1.f);
    ConvergeChecker criteria(config.finthres);
    std::vector<double> objlist;
    DevField c(ne), rhat(ne), x(ne), sens(ne), t(ne), t2(ne), objGrad(ne), gGrad(ne), volGrad(ne);
    // dJ/drho for dJ/dC = dC through x = rhat^3, rhat = sgm(c), c = conv(rho)
    auto backward = [&](float dC[6][6], DevField& grad) {
        hom.Sensitivity(dC, sens.p, hom.grid->cellReso[0], true);
        ops.pow_grad(rhat.p, sens.p, t.p, 3.0);
        ops.sgm(t.p, rhat.p, t2.p);
        ops.spline4_filter(t2.p, grad.p, config.filterRadius, true);
    };
    double C[6][6];
    for (int iter = 0; iter < config.max_iter; iter++) {
        ops.spline4_filter(rho.p, c.p, config.filterRadius, false);
        ops.sgm(c.p, nullptr, rhat.p);
        ops.pow(rhat.p, x.p, 3.0);
        hom.update(x.p, hom.grid->cellReso[0]);
        hom.elasticMatrix(C);
        float dC[6][6] = {};
        const double val = shear_objective(C, dC);
        objlist.push_back(val);
        backward(dC, objGrad);
        std::printf(" * Iter %d   obj = %.4e   (solve %.3fs)\n", iter, val, hom.last_stats.seconds);
        if (criteria.is_converge(iter, val)) {
            std::printf("= converged\n");
            break;
        }
        ops.symmetrize(objGrad.p, config.sym, config.usesym);
        MmaConstraint k = shear_isotropy_constraint(C);
        backward(k.dC, gGrad);
        ops.symmetrize(gGrad.p, config.sym, config.usesym);
        const double vol_ratio = ops.sum(rhat.p) / ne;
        float gval[2] = {(float)((vol_ratio - config.volRatio) * 1000.0), (float)k.value};
        ops.fill(t.p, 1.f / ne);
        ops.sgm(t.p, rhat.p, t2.p);
        ops.spline4_filter(t2.p, volGrad.p, config.filterRadius, true);
        std::printf("zener ratio = %4.2e ; obj = %4.2e ; vol = %4.2e\n", std::sqrt(std::max(k.value + 0.1, 0.0) / 1e6) + 1,
                    val, vol_ratio);
        float* dgdx[2] = {volGrad.p, gGrad.p};
        mma.update(iter, rho.p, objGrad.p, gval, dgdx);
        logIter(iter, config, rho, C, val);
    }
    {
        const std::vector<float> rh = rho.download();
        std::ofstream(out_path(config, "rho"), std::ios::binary)
            .write(reinterpret_cast<const char*>(rh.data()), sizeof(float) * rh.size());
    }
    float Cf[36];
    for (int i = 0; i < 36; i++) Cf[i] = (float)C[i / 6][i % 6];
    std::ofstream(out_path(config, "C"), std::ios::binary).write(reinterpret_cast<const char*>(Cf), sizeof(Cf));
    std::ofstream ol(out_path(config, "objlist"));
    for (double v : objlist) ol << v << "\n";
    return objlist;
}

// Negative Poisson ratio by MMA with a volume constraint (the 256^3 MMA workload): the objective of
// optiNpr2 (testAutoDiff.cu:466-493) plus one volume constraint.
std::vector<double> example_opti_npr_mma(cfg::HomoConfig config) {
    auto obj = [](int iter, const double C[6][6], float dC[6][6]) {
        const double beta = std::pow(0.8, iter);
        dC[0][1] = dC[0][2] = dC[1][2] = 1.f;
        dC[0][0] = dC[1][1] = dC[2][2] = (float)-beta;
        return C[0][1] + C[0][2] + C[1][2] - (C[0][0] + C[1][1] + C[2][2]) * beta;
    };
    return run_mma(config, obj, [](const double (*)[6]) { return std::vector<MmaConstraint>{}; });
}

// user hook (Framework.cu:359-361)
__attribute__((weak)) void example_yours(cfg::HomoConfig) {}

// runCustom (Framework.cu:362-366)
void runCustom(cfg::HomoConfig config) { example_yours(config); }

// runInstance (homogenization/testAutoDiff.cu:571-607): dispatch on the objective and optimiser.
std::vector<double> runInstance(cfg::HomoConfig config) {
    switch (config.obj) {
        case cfg::Objective::bulk: return example_opti_bulk(config);
        case cfg::Objective::npr: return config.optimizer == "mma" ? example_opti_npr_mma(config) : example_opti_npr(config);
        case cfg::Objective::shear: return example_opti_shear_isotropy(config);
        case cfg::Objective::custom: runCustom(config); return {};
    }
    return {};
}

}  // namespace homo

// main (main.cpp:22-48) as a C entry point: parse flags, run the instance, map exceptions to exit codes.
// The shear-isotropy objective and constraint terms for a given C (row-major double[36]), for callers
// that assemble the optimisation themselves: obj, dJ/dC, constraint value and dg/dC (float[36] each).
extern "C" int homo3d_shear_isotropy_terms(const double* C36, double* obj, float* dobj36, double* g, float* dg36) {
    if (!C36 || !obj || !dobj36 || !g || !dg36) return 3;
    double C[6][6];
    for (int i = 0; i < 36; i++) C[i / 6][i % 6] = C36[i];
    float dC[6][6] = {};
    *obj = homo::shear_objective(C, dC);
    for (int i = 0; i < 36; i++) dobj36[i] = dC[i / 6][i % 6];
    const auto k = homo::shear_isotropy_constraint(C);
    *g = k.value;
    for (int i = 0; i < 36; i++) dg36[i] = k.dC[i / 6][i % 6];
    return 0;
}

extern "C" int homo3d_framework_main(int argc, char** argv) {
    try {
        cfg::HomoConfig config;
        config.parse(argc, argv);
        auto objs = homo::runInstance(config);
        std::printf("finished %zu iterations\n", objs.size());
        return 0;
    } catch (const homo::Homo3dError& e) {
        std::fprintf(stderr, "homo3d error %d: %s\n", e.code(), e.what());
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
