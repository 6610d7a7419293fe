// bemhip.hpp -- header-only C++ surface over the C-ABI in bemhip.h that keeps
// the reference's Bem::ColocSim call surface (ColocSim.hpp:16-22,
// LinLinSim.hpp:22-67, Simulation.hpp:29-167): same method names, argument
// meaning and error behaviour (std::out_of_range("negative values") from the
// wall-mode RK4 step, LinLinSim.cpp:157).  Everything runs on the GPU; state
// stays resident between calls.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "bemhip.h"

namespace bemhip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

class Context {
public:
    explicit Context(int device = 0) {
        const int rc = bemhip_create(device, &h_);
        if (rc != BEMHIP_OK) throw Error(rc, "bemhip_create failed");
    }
    ~Context() { bemhip_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    bemhip_ctx* get() const { return h_; }
    void check(int rc) const {
        if (rc == BEMHIP_ERR_NEG_X_IN_MIRROR) throw std::out_of_range("negative values");
        if (rc != BEMHIP_OK) throw Error(rc, bemhip_last_error(h_));
    }

private:
    bemhip_ctx* h_ = nullptr;
};

using vec3d = std::array<double, 3>;
using tri3 = std::array<std::size_t, 3>;

// p_field(x, t) = amp * sin(2 pi freq t + phase); the reference takes a
// function pointer (Simulation.hpp:29), which cannot run on the device.
struct PressureField {
    double amp = 0.0, freq = 0.0, phase = 0.0;
};

class ColocSim {
public:
    ColocSim(Context& ctx, const std::vector<vec3d>& verts, const std::vector<tri3>& trigs, double p_inf = 1.0,
             double epsilon = 1.0, double sigma = 0.0, double gamma = 1.0, PressureField field = {},
             bool mirror = false, bool independent_parts = false)
        : ctx_(ctx) {
        bemhip_sim_default_params(&p_);
        p_.p_inf = p_inf; p_.epsilon = epsilon; p_.sigma = sigma; p_.gamma = gamma;
        p_.field_amp = field.amp; p_.field_freq = field.freq; p_.field_phase = field.phase;
        p_.flags = (mirror ? BEMHIP_MIRROR_X0 : 0u) | (independent_parts ? BEMHIP_INDEPENDENT_PARTS : 0u);
        p_.wall_guards = mirror ? 1 : 0;
        std::vector<uint32_t> t = narrow(trigs, verts.size());
        ctx_.check(bemhip_sim_create(ctx_.get(), &p_, (int64_t)verts.size(), verts.empty() ? nullptr : verts[0].data(),
                                     (int64_t)trigs.size(), t.data(), &h_));
        N_ = verts.size();
    }
    ~ColocSim() { bemhip_sim_destroy(h_); }
    ColocSim(const ColocSim&) = delete;
    ColocSim& operator=(const ColocSim&) = delete;

    // LinLinSim::evolve_system_RK4 / evolve_system (LinLinSim.cpp:194-284)
    double evolve_system_RK4(double dp, bool fixdt = false) {
        double dt = 0.0;
        ctx_.check(bemhip_sim_evolve_rk4(h_, dp, fixdt ? 1 : 0, &dt));
        return dt;
    }
    double evolve_system(double dp, bool fixdt = false) {
        double dt = 0.0;
        ctx_.check(bemhip_sim_evolve_euler(h_, dp, fixdt ? 1 : 0, &dt));
        return dt;
    }
    void compute_psi() { ctx_.check(bemhip_sim_compute_psi(h_)); }
    std::vector<vec3d> position_t() {
        std::vector<vec3d> v(N_);
        ctx_.check(bemhip_sim_position_t(h_, v[0].data(), nullptr));
        return v;
    }
    std::vector<double> pot_t_multi(const std::vector<vec3d>& grads, double t) {
        std::vector<double> out(N_);
        ctx_.check(bemhip_sim_pot_t_multi(h_, grads[0].data(), t, out.data()));
        return out;
    }

    // Simulation setters / getters (Simulation.hpp:92-167)
    void set_phi(const std::vector<double>& v) { ctx_.check(bemhip_sim_set_phi(h_, v.data())); }
    void set_phi(double value) { set_phi(std::vector<double>(N_, value)); }
    void set_psi(const std::vector<double>& v) { ctx_.check(bemhip_sim_set_psi(h_, v.data())); }
    void set_V_0(double v) { ctx_.check(bemhip_sim_set_V_0(h_, v)); }
    void set_time(double t) { ctx_.check(bemhip_sim_set_time(h_, t)); }
    void set_min_dt(double v) { p_.min_dt = v; push(); }
    void set_dp_balance(double v) { p_.dp_balance = v; push(); }
    void set_bcgstab(bool v) { p_.solver = v ? BEMHIP_SOLVE_BICGSTAB64 : BEMHIP_SOLVE_LU64; push(); }
    void set_num_threads(std::size_t) {}   // OpenMP thread count of the CPU path; no meaning here
    void set_solver(bemhip_solver s, double tol = 0.0, int maxit = 0) {
        p_.solver = s; p_.tol = tol; p_.maxit = maxit; push();
    }
    void set_mesh(const std::vector<vec3d>& verts, const std::vector<tri3>& trigs) {
        std::vector<uint32_t> t = narrow(trigs, verts.size());
        ctx_.check(bemhip_sim_set_mesh(h_, (int64_t)verts.size(), verts[0].data(), (int64_t)trigs.size(), t.data()));
        N_ = verts.size();
    }
    double get_time() const { double t; state(nullptr, nullptr, nullptr, &t); return t; }
    double get_volume() const { double v; ctx_.check(bemhip_sim_get_volume(h_, &v)); return v; }
    // loose parts and, for a batch of independent bubbles, their own V_0 and forcing
    int nparts() const { int32_t n = 0; ctx_.check(bemhip_sim_get_parts(h_, &n)); return n; }
    std::vector<double> part_volumes() const {
        std::vector<double> v(nparts());
        ctx_.check(bemhip_sim_get_part_volumes(h_, v.data(), (int32_t)v.size()));
        return v;
    }
    void set_part_physics(const std::vector<double>& V_0, const std::vector<double>& amp = {},
                          const std::vector<double>& freq = {}, const std::vector<double>& phase = {}) {
        auto ptr = [](const std::vector<double>& a) { return a.empty() ? nullptr : a.data(); };
        ctx_.check(bemhip_sim_set_part_physics(h_, nparts(), ptr(V_0), ptr(amp), ptr(freq), ptr(phase)));
    }
    std::vector<vec3d> get_vertices() const { std::vector<vec3d> v(N_); state(v[0].data(), nullptr, nullptr, nullptr); return v; }
    std::vector<double> get_phi() const { std::vector<double> v(N_); state(nullptr, v.data(), nullptr, nullptr); return v; }
    std::vector<double> get_psi() const { std::vector<double> v(N_); state(nullptr, nullptr, v.data(), nullptr); return v; }
    bemhip_solve_stats last_stats() const { bemhip_solve_stats s; ctx_.check(bemhip_sim_last_stats(h_, &s)); return s; }

    // Loss sum_n (R_n - target_n)^2 of the volume-equivalent radius over
    // target.size() fixed RK4 steps and its gradient with respect to the
    // current vertices and potential (state left unchanged); see
    // bemhip_sim_adjoint_radius.
    double adjoint_radius(double dt, const std::vector<double>& target, int every, std::vector<vec3d>* grad_x,
                          std::vector<double>* grad_phi, std::vector<double>* radius = nullptr) {
        double loss = 0.0;
        if (grad_x) grad_x->assign(N_, vec3d{0.0, 0.0, 0.0});
        if (grad_phi) grad_phi->assign(N_, 0.0);
        if (radius) radius->assign(target.size(), 0.0);
        ctx_.check(bemhip_sim_adjoint_radius(h_, (int32_t)target.size(), dt, target.data(), every, &loss,
                                             radius ? radius->data() : nullptr,
                                             grad_x ? (*grad_x)[0].data() : nullptr,
                                             grad_phi ? grad_phi->data() : nullptr));
        return loss;
    }

private:
    static std::vector<uint32_t> narrow(const std::vector<tri3>& trigs, std::size_t n) {
        std::vector<uint32_t> t;
        t.reserve(3 * trigs.size());
        for (const tri3& tr : trigs)
            for (std::size_t k : tr) {
                if (k >= n || k > 0xFFFFFFFFu) throw std::out_of_range("out of range");   // Triplet.hpp:33
                t.push_back((uint32_t)k);
            }
        return t;
    }
    void push() {
        bemhip_sim_params q = p_;
        q.V_0 = std::numeric_limits<double>::quiet_NaN();   // keep the simulation's V_0
        ctx_.check(bemhip_sim_set_params(h_, &q));
    }
    void state(double* x, double* phi, double* psi, double* t) const {
        ctx_.check(bemhip_sim_get_state(h_, x, phi, psi, t));
    }
    Context& ctx_;
    bemhip_sim* h_ = nullptr;
    bemhip_sim_params p_{};
    std::size_t N_ = 0;
};

}  // namespace bemhip
