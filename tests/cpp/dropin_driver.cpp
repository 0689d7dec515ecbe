// Runs the reference's simulation() shape through the C++ drop-in adapter:
// load params and the initial state, call hip_time_integration(...) where
// the reference calls time_integration(...) (sources/SPH.cpp:63-68), catch
// std::exception and print it as sources/SPH.cpp:76-85 does.
//
//   dropin_driver <params.txt> <state.bin> <out.bin> [fp32] [probe_every]
//
// params.txt : one "name value" per line (the ExperimentParams field names);
//              piston_a / piston_b take a comma-separated list
// state.bin  : uint32 n, then r[dim n], v[dim n], rho[n] as float64, itype[n] as int32
// out.bin    : uint64 steps, float64 time, float64 dt, then r, v, rho, p (float64)
//              and itype (int32) of the final state, then uint64 probe_step and
//              the rho and p that the output loaders returned at that step
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rrsph_hip.hpp"
#include "reference_shapes.hpp"

namespace {

// The fields of ExperimentParams the adapter reads, under the reference's names.
struct Params {
    unsigned dim = 2, maxn = 0, ntotal = 0, nfluid = 0, max_neighbours = 64;
    double x_mingeom = 0, x_maxgeom = 0, y_mingeom = 0, y_maxgeom = 0, z_mingeom = 0, z_maxgeom = 0;
    double hsml = 0, delta = 0, mass = 0, rho0 = 1000, cell_scale_k = 2, depth = 0, g = 9.81,
           pi = 3.14159265358979323846;
    unsigned eos_sound_vel_method = 0;
    double eos_sound_vel_coef = 1, eos_sound_vel = 0;
    unsigned density_treatment = 1, density_skf = 1, density_normalization = 0;
    double density_delta_sph_coef = 0.1;
    unsigned intf_sph_approximation = 2, intf_skf = 1;
    bool artificial_pressure = false;
    unsigned artificial_pressure_skf = 1;
    double artificial_pressure_index = 4, artificial_pressure_coef = 0.2;
    bool artificial_viscosity = true;
    unsigned artificial_viscosity_skf = 1;
    double artificial_shear_visc = 1, artificial_bulk_visc = 0;
    bool average_velocity = true;
    unsigned average_velocity_skf = 1;
    double average_velocity_coef = 0.3;
    unsigned boundary_treatment = 0, dt_correction_method = 0;
    double dt = 0, CFL_coef = 0;
    bool consistency_check = true, use_crash_dump = true;
    unsigned consistency_treatment = 0;
    unsigned nwm = 0, nwm_direction = 0, nwm_particles_start = 0, nwm_particles_end = 0;
    double nwm_time_start = 0, nwm_freq = 0, nwm_wave_number = 0, nwm_wave_magnitude = 0, nwm_phase = 0;
    double start_simulation_time = 0, simulation_time = 0;
};

std::vector<double> parse_list(const std::string& s) {
    std::vector<double> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ','))
        if (!item.empty()) out.push_back(std::stod(item));
    return out;
}

void load_params(const char* path, Params& p, rrsph_hip::Options& o) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error(std::string("can't open ") + path);
    std::map<std::string, std::string> kv;
    std::string k, v;
    while (in >> k >> v) kv[k] = v;
    auto num = [&](const char* name, auto& field) {
        auto it = kv.find(name);
        if (it != kv.end()) field = static_cast<std::decay_t<decltype(field)>>(std::stod(it->second));
    };
#define F(name) num(#name, p.name)
    F(dim); F(maxn); F(ntotal); F(nfluid); F(max_neighbours);
    F(x_mingeom); F(x_maxgeom); F(y_mingeom); F(y_maxgeom); F(z_mingeom); F(z_maxgeom);
    F(hsml); F(delta); F(mass); F(rho0); F(cell_scale_k); F(depth); F(g); F(pi);
    F(eos_sound_vel_method); F(eos_sound_vel_coef); F(eos_sound_vel);
    F(density_treatment); F(density_skf); F(density_delta_sph_coef);
    F(intf_sph_approximation); F(intf_skf);
    F(artificial_pressure); F(artificial_pressure_skf); F(artificial_pressure_index); F(artificial_pressure_coef);
    F(artificial_viscosity); F(artificial_viscosity_skf); F(artificial_shear_visc); F(artificial_bulk_visc);
    F(average_velocity); F(average_velocity_skf); F(average_velocity_coef);
    F(boundary_treatment); F(dt_correction_method); F(dt); F(CFL_coef);
    F(consistency_check); F(consistency_treatment); F(use_crash_dump);
    F(nwm); F(nwm_direction); F(nwm_particles_start); F(nwm_particles_end);
    F(nwm_time_start); F(nwm_freq); F(nwm_wave_number); F(nwm_wave_magnitude); F(nwm_phase);
    F(start_simulation_time); F(simulation_time);
#undef F
    bool motion = false;
    num("boundary_motion", motion);
    o.boundary_motion = motion;
    num("dt_dynamic_mode", o.dt_dynamic_mode);
    num("piston_omega", o.piston_omega);
    if (kv.count("piston_a")) o.piston_a = parse_list(kv["piston_a"]);
    if (kv.count("piston_b")) o.piston_b = parse_list(kv["piston_b"]);
}

template <class T> void read_raw(std::ifstream& in, T* dst, size_t n) {
    in.read(reinterpret_cast<char*>(dst), n * sizeof(T));
    if (!in) throw std::runtime_error("state file too short");
}
template <class T> void write_raw(std::ofstream& out, const T* src, size_t n) {
    out.write(reinterpret_cast<const char*>(src), n * sizeof(T));
}

int run(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "usage: dropin_driver <params.txt> <state.bin> <out.bin> [fp32] [probe_every]\n";
        return 2;
    }
    Params params;
    rrsph_hip::Options opt;
    load_params(argv[1], params, opt);
    if (argc > 4 && std::string(argv[4]) == "fp32") opt.precision = RRSPH_FP32;
    shapes::Output output;
    if (argc > 5) output.probe_every = std::stoul(argv[5]);

    std::ifstream in(argv[2], std::ios::binary);
    if (!in) throw std::runtime_error(std::string("can't open ") + argv[2]);
    uint32_t n = 0;
    read_raw(in, &n, 1);
    if (n != params.ntotal) throw std::runtime_error("state size differs from ntotal");
    if (params.maxn < n) params.maxn = n;
    const int dim = (int)params.dim;
    std::vector<double> r(dim * n), v(dim * n), rho(n);
    std::vector<int32_t> it(n);
    read_raw(in, r.data(), r.size());
    read_raw(in, v.data(), v.size());
    read_raw(in, rho.data(), rho.size());
    read_raw(in, it.data(), it.size());

    shapes::vheap_darray_floatn r_var(params.maxn), v_var(params.maxn);
    shapes::heap_darray<double> rho_a(params.maxn), p_a(params.maxn);
    shapes::heap_darray<int> itype(params.maxn);
    rrsph_hip::aos_to_floatn(r, n, dim, r_var);
    rrsph_hip::aos_to_floatn(v, n, dim, v_var);
    for (uint32_t i = 0; i < n; ++i) {
        rho_a(i) = rho[i];
        itype(i) = it[i];
    }

    try {
        rrsph_hip::hip_time_integration(params, output, r_var, v_var, rho_a, p_a, itype, opt);
    } catch (...) {
        if (output.crashed)
            std::cout << "crash dump at step " << output.crash_step << " time " << output.crash_time << std::endl;
        throw;
    }

    std::ofstream out(argv[3], std::ios::binary);
    const uint64_t steps = output.last_step;
    const double t = output.times.empty() ? params.start_simulation_time : output.times.back();
    write_raw(out, &steps, 1);
    write_raw(out, &t, 1);
    write_raw(out, &params.dt, 1);
    rrsph_hip::floatn_to_aos(r_var, n, dim, r);
    rrsph_hip::floatn_to_aos(v_var, n, dim, v);
    write_raw(out, r.data(), r.size());
    write_raw(out, v.data(), v.size());
    for (uint32_t i = 0; i < n; ++i) write_raw(out, &rho_a(i), 1);
    for (uint32_t i = 0; i < n; ++i) write_raw(out, &p_a(i), 1);
    for (uint32_t i = 0; i < n; ++i) {
        int32_t x = itype(i);
        write_raw(out, &x, 1);
    }
    const uint64_t ps = output.probe_step;
    write_raw(out, &ps, 1);
    for (uint32_t i = 0; i < n; ++i) {
        double x = output.probe_rho ? (*output.probe_rho)(i) : 0.0;
        write_raw(out, &x, 1);
    }
    for (uint32_t i = 0; i < n; ++i) {
        double x = output.probe_p ? (*output.probe_p)(i) : 0.0;
        write_raw(out, &x, 1);
    }
    std::cout << "steps " << steps << " time " << t << " dt " << params.dt << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& ex) {
        std::cout << "Catch exception: " << ex.what() << std::endl;
        return 1;
    }
}
