// RRSPH_HIP -- C++ drop-in for the reference's backend entry point.
//
// Header-only adapter over the C-ABI of rrsph_hip.h.  It is meant to be
// compiled inside the reference's own build (SPH.cpp, `#elif defined
// RRSPH_HIP`), so it is templated on the reference's containers and global
// params instead of naming them:
//
//   hip_time_integration(params, output, r_var, v_var, rho, p, itype)
//     replaces time_integration(r_var, v_var, rho, p, itype)
//     (sources/RRSPH_OMP/include/TimeIntegration.h:4-9, src/TimeIntegration.cpp:21-103)
//     and cl_time_integration (sources/RRSPH_CL/include/CLAdapter.h:10-15).
//
// Requirements on the template arguments (all met by the reference's types):
//   Params      : the fields of ExperimentParams (sources/RRSPH_PARAMS/include/Params.h:8-194)
//   Output      : setup_output(load_r, load_itype, load_v, load_p, load_rho),
//                 start_step(t), update_step(t, step), finish_step(),
//                 crash_dump(r, itype, v, rho, p, step, time)
//                 (sources/RRSPH_COMMON/include/Output.h:11-56)
//   FloatnArray : vheap_darray_floatn -- get_flt2() / get_flt3() returning an
//                 indexable array of {x, y} / {x, y, z}; constructible from a size
//   FloatArray / IntArray : heap_darray<rr_float> / heap_darray<rr_int>
// Errors are thrown as std::runtime_error with the reference's messages.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rrsph_hip.h"

namespace rrsph_hip {

class Error : public std::runtime_error {
public:
    Error(rrsph_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    rrsph_status status;
};

struct Options {
    rrsph_precision precision = RRSPH_FP64;
    // RRSPH_CL behaviour: boundaries move, wave maker active
    bool boundary_motion = false;
    // DT_CORRECTION_DYNAMIC: 0 = OpenMP build of RRSPH_OMP, 1 = max over particles
    int32_t dt_dynamic_mode = 0;
    // nwm = 6 (Fourier piston, an extension of this library): piston
    // velocity sum_k a_k cos(k w t) + b_k sin(k w t), at most 32 harmonics
    double piston_omega = 0;
    std::vector<double> piston_a, piston_b;
    // density_normalization (summation density): renormalise every this many steps
    uint32_t density_normalization_every = 1;
};

// c^2 and c of the EOS as the OpenMP backend computes them (eos_sound_vel_method 1
// takes eos_sound_vel itself as c^2, sources/RRSPH_COMMON/src/EOS.cpp:3-20)
template <class Params> inline void eos_constants(const Params& p, double& c_sqr, double& c) {
    c_sqr = p.eos_sound_vel_method == 0 ? 200 * p.g * p.depth * p.eos_sound_vel_coef : (double)p.eos_sound_vel;
    c = std::sqrt(c_sqr);
}

// Flatten the reference's params (after fileInput) into the C-ABI struct.
template <class Params> inline rrsph_params_c make_params(const Params& p, const Options& o = Options()) {
    rrsph_params_c c;
    std::memset(&c, 0, sizeof(c));
    c.struct_size = sizeof(rrsph_params_c);
    c.dim = (int32_t)p.dim;
    c.ntotal = (uint32_t)p.ntotal;
    c.nfluid = (uint32_t)p.nfluid;
    c.max_neighbours = (uint32_t)p.max_neighbours;
    c.x_mingeom = p.x_mingeom;
    c.x_maxgeom = p.x_maxgeom;
    c.y_mingeom = p.y_mingeom;
    c.y_maxgeom = p.y_maxgeom;
    c.z_mingeom = p.z_mingeom;
    c.z_maxgeom = p.z_maxgeom;
    c.hsml = p.hsml;
    c.delta = p.delta;
    c.mass = p.mass;
    c.rho0 = p.rho0;
    c.cell_scale_k = p.cell_scale_k;
    c.depth = p.depth;
    c.g = p.g;
    c.pi = p.pi;
    eos_constants(p, c.eos_c_sqr, c.eos_c);
    c.eos_sound_vel = p.eos_sound_vel;
    c.density_treatment = (int32_t)p.density_treatment;
    c.density_skf = (int32_t)p.density_skf;
    c.density_delta_sph_coef = p.density_delta_sph_coef;
    c.density_normalization = (int32_t)p.density_normalization;
    c.density_normalization_every = o.density_normalization_every;
    c.intf_sph_approximation = (int32_t)p.intf_sph_approximation;
    c.intf_skf = (int32_t)p.intf_skf;
    c.artificial_pressure = p.artificial_pressure ? 1 : 0;
    c.artificial_pressure_skf = (int32_t)p.artificial_pressure_skf;
    c.artificial_pressure_index = p.artificial_pressure_index;
    c.artificial_pressure_coef = p.artificial_pressure_coef;
    c.artificial_viscosity = p.artificial_viscosity ? 1 : 0;
    c.artificial_viscosity_skf = (int32_t)p.artificial_viscosity_skf;
    c.artificial_shear_visc = p.artificial_shear_visc;
    c.artificial_bulk_visc = p.artificial_bulk_visc;
    c.average_velocity = p.average_velocity ? 1 : 0;
    c.average_velocity_skf = (int32_t)p.average_velocity_skf;
    c.average_velocity_coef = p.average_velocity_coef;
    c.boundary_treatment = (int32_t)p.boundary_treatment;
    c.dt_correction_method = (int32_t)p.dt_correction_method;
    c.dt = p.dt;
    c.CFL_coef = p.CFL_coef;
    c.consistency_check = p.consistency_check ? 1 : 0;
    c.consistency_treatment = (int32_t)p.consistency_treatment;
    c.nwm = (int32_t)p.nwm;
    c.nwm_direction = (int32_t)p.nwm_direction;
    c.nwm_particles_start = (uint32_t)p.nwm_particles_start;
    c.nwm_particles_end = (uint32_t)p.nwm_particles_end;
    c.nwm_time_start = p.nwm_time_start;
    c.nwm_freq = p.nwm_freq;
    c.nwm_wave_number = p.nwm_wave_number;
    c.nwm_wave_magnitude = p.nwm_wave_magnitude;
    c.nwm_phase = p.nwm_phase;
    c.boundary_motion = o.boundary_motion ? 1 : 0;
    c.dt_dynamic_mode = o.dt_dynamic_mode;
    c.piston_nharm = (int32_t)std::min<size_t>(std::min(o.piston_a.size(), o.piston_b.size()), 32);
    c.piston_omega = o.piston_omega;
    for (int32_t k = 0; k < c.piston_nharm; ++k) {
        c.piston_a[k] = o.piston_a[k];
        c.piston_b[k] = o.piston_b[k];
    }
    c.precision = (int32_t)o.precision;
    c.start_time = p.start_simulation_time;
    return c;
}

// RAII owner of one RRSPH_HIP context.
class Context {
public:
    explicit Context(const rrsph_params_c& prm) {
        rrsph_status st = rrsph_hip_create(&prm, &ctx_);
        if (st != RRSPH_OK) {
            std::string msg = ctx_ ? rrsph_hip_last_error(ctx_) : "rrsph_hip_create failed";
            if (ctx_) rrsph_hip_destroy(ctx_);
            ctx_ = nullptr;
            throw Error(st, msg);
        }
        n_ = prm.ntotal;
    }
    ~Context() {
        if (ctx_) rrsph_hip_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    void check(rrsph_status st) const {
        if (st != RRSPH_OK) throw Error(st, rrsph_hip_last_error(ctx_));
    }
    void upload(const double* r, const double* v, const double* rho, const int32_t* itype) {
        check(rrsph_hip_upload(ctx_, r, v, rho, itype, n_));
    }
    void step(uint32_t n = 1) { check(rrsph_hip_step(ctx_, n)); }
    void time(double& t, double& dt, uint64_t& step) const { check(rrsph_hip_get_time(ctx_, &t, &dt, &step)); }
    void download(uint32_t fields, double* r, double* v, double* rho, double* p, int32_t* itype) const {
        check(rrsph_hip_download(ctx_, fields, r, v, rho, p, itype));
    }
    uint32_t size() const { return n_; }
    rrsph_ctx* get() const { return ctx_; }

private:
    rrsph_ctx* ctx_ = nullptr;
    uint32_t n_ = 0;
};

// Copy between the reference's vector arrays and AoS doubles (dim per particle).
template <class FloatnArray> void floatn_to_aos(FloatnArray& a, size_t n, int dim, std::vector<double>& out) {
    out.resize(dim * n);
    if (dim == 3) {
        auto& f = a.get_flt3();
        for (size_t i = 0; i < n; ++i) {
            out[3 * i] = f(i).x;
            out[3 * i + 1] = f(i).y;
            out[3 * i + 2] = f(i).z;
        }
    } else {
        auto& f = a.get_flt2();
        for (size_t i = 0; i < n; ++i) {
            out[2 * i] = f(i).x;
            out[2 * i + 1] = f(i).y;
        }
    }
}
template <class FloatnArray> void aos_to_floatn(const std::vector<double>& in, size_t n, int dim, FloatnArray& a) {
    if (dim == 3) {
        auto& f = a.get_flt3();
        for (size_t i = 0; i < n; ++i) {
            f(i).x = in[3 * i];
            f(i).y = in[3 * i + 1];
            f(i).z = in[3 * i + 2];
        }
    } else {
        auto& f = a.get_flt2();
        for (size_t i = 0; i < n; ++i) {
            f(i).x = in[2 * i];
            f(i).y = in[2 * i + 1];
        }
    }
}

// Drop-in for time_integration(r_var, v_var, rho, p, itype): runs the
// reference's loop `while (time <= simulation_time)` on the GPU, keeps
// params.dt current (the reference mutates it inside the step) and serves
// the RRSPHOutput loaders from device memory in original particle order.
template <class Params, class Output, class FloatnArray, class FloatArray, class IntArray>
void hip_time_integration(Params& params, Output& output, FloatnArray& r_var, FloatnArray& v_var, FloatArray& rho,
                          FloatArray& p, IntArray& itype, const Options& opt = Options()) {
    const size_t n = params.ntotal;
    const int dim = (int)params.dim;
    std::vector<double> hr, hv, hrho(n);
    std::vector<int32_t> hit(n);
    Context ctx(make_params(params, opt));
    floatn_to_aos(r_var, n, dim, hr);
    floatn_to_aos(v_var, n, dim, hv);
    for (size_t i = 0; i < n; ++i) {
        hrho[i] = rho(i);
        hit[i] = (int32_t)itype(i);
    }
    ctx.upload(hr.data(), hv.data(), hrho.data(), hit.data());

    const size_t maxn = params.maxn;
    auto load_floatn = [&ctx, n, maxn, dim](bool vel) {
        std::vector<double> buf(dim * n);
        if (vel) ctx.download(RRSPH_FIELD_V, nullptr, buf.data(), nullptr, nullptr, nullptr);
        else ctx.download(RRSPH_FIELD_R, buf.data(), nullptr, nullptr, nullptr, nullptr);
        auto out = std::make_shared<FloatnArray>(maxn);
        aos_to_floatn(buf, n, dim, *out);
        return out;
    };
    auto load_scalar = [&ctx, n, maxn](uint32_t field) {
        std::vector<double> buf(n);
        ctx.download(field, nullptr, nullptr, field == RRSPH_FIELD_RHO ? buf.data() : nullptr,
                     field == RRSPH_FIELD_P ? buf.data() : nullptr, nullptr);
        auto out = std::make_shared<FloatArray>(maxn);
        for (size_t i = 0; i < n; ++i) (*out)(i) = buf[i];
        return out;
    };
    auto load_itype = [&ctx, n, maxn]() {
        std::vector<int32_t> buf(n);
        ctx.download(RRSPH_FIELD_ITYPE, nullptr, nullptr, nullptr, nullptr, buf.data());
        auto out = std::make_shared<IntArray>(maxn);
        for (size_t i = 0; i < n; ++i) (*out)(i) = buf[i];
        return out;
    };
    output.setup_output([=]() { return load_floatn(false); }, [=]() { return load_itype(); },
                        [=]() { return load_floatn(true); }, [=]() { return load_scalar(RRSPH_FIELD_P); },
                        [=]() { return load_scalar(RRSPH_FIELD_RHO); });

    // The in-bounds check of every step (RRSPHOutput::update_step,
    // Output.cpp:366-412) runs on the device: the library halts with the
    // reference's error under STOP.  update_step therefore runs with the
    // host check off (it would download r and itype every step), and the
    // crash dump the reference writes before rethrowing is written here.
    const bool check = params.consistency_check;
    double time = params.start_simulation_time, dt = params.dt;
    uint64_t step = 0;
    while (time <= params.simulation_time) {
        output.start_step(time);
        try {
            ctx.step(1);
        } catch (const Error& e) {
            if (e.status == RRSPH_ERR_OUTSIDE_BOUNDARIES && params.use_crash_dump) {
                ctx.time(time, dt, step);
                output.crash_dump(load_floatn(false), load_itype(), load_floatn(true), load_scalar(RRSPH_FIELD_RHO),
                                  load_scalar(RRSPH_FIELD_P), step, time);
            }
            throw;
        }
        ctx.time(time, dt, step);
        params.dt = dt;
        params.consistency_check = false;
        output.update_step(time, step);
        params.consistency_check = check;
        output.finish_step();
    }
    // leave the final state in the caller's arrays, as RRSPH_OMP does
    std::vector<double> hp(n);
    ctx.download(RRSPH_FIELD_R | RRSPH_FIELD_V | RRSPH_FIELD_RHO | RRSPH_FIELD_P | RRSPH_FIELD_ITYPE, hr.data(),
                 hv.data(), hrho.data(), hp.data(), hit.data());
    aos_to_floatn(hr, n, dim, r_var);
    aos_to_floatn(hv, n, dim, v_var);
    for (size_t i = 0; i < n; ++i) {
        rho(i) = hrho[i];
        p(i) = hp[i];
        itype(i) = hit[i];
    }
}

}  // namespace rrsph_hip
