// Minimal containers with the interface shape the reference's backend entry
// point uses (our own code, written from the interface alone):
//   heap_darray<T>          -- operator()(i), size()
//                              (sources/RR/include/RR/Memory/HeapDArray.h:130-235)
//   vheap_darray_floatn     -- get_flt2() / get_flt3() -> heap_darray<rr_float2 / 3>
//                              (sources/RRSPH_PARAMS/include/RRFloatnDArray.h:7-44)
//   Output                  -- setup_output(load_r, load_itype, load_v, load_p,
//                              load_rho), start_step, update_step, finish_step
//                              (sources/RRSPH_COMMON/include/Output.h:11-31)
// They let the drop-in adapter of include/rrsph_hip.hpp be compiled and run
// without the reference's build.
#pragma once

#include <cstddef>
#include <functional>
#include <memory>
#include <vector>

namespace shapes {

template <class T> class heap_darray {
public:
    heap_darray() = default;
    explicit heap_darray(size_t n) : v_(n) {}
    T& operator()(size_t i) { return v_[i]; }
    const T& operator()(size_t i) const { return v_[i]; }
    size_t size() const { return v_.size(); }

private:
    std::vector<T> v_;
};

struct rr_float2 {
    double x = 0, y = 0;
};
struct rr_float3 {
    double x = 0, y = 0, z = 0;
};

// holds both shapes (the reference keeps one of them in a std::variant)
class vheap_darray_floatn {
public:
    vheap_darray_floatn() = default;
    explicit vheap_darray_floatn(size_t n) : a_(n), b_(n) {}
    heap_darray<rr_float2>& get_flt2() { return a_; }
    const heap_darray<rr_float2>& get_flt2() const { return a_; }
    heap_darray<rr_float3>& get_flt3() { return b_; }
    const heap_darray<rr_float3>& get_flt3() const { return b_; }

private:
    heap_darray<rr_float2> a_;
    heap_darray<rr_float3> b_;
};

using shared_floatn = std::shared_ptr<vheap_darray_floatn>;
template <class T> using shared_darray = std::shared_ptr<heap_darray<T>>;

// Records the step times and, every `probe_every` steps, calls every loader
// (as the reference's output thread does when it saves a frame).
class Output {
public:
    using load_floatn = std::function<shared_floatn()>;
    using load_float = std::function<shared_darray<double>()>;
    using load_int = std::function<shared_darray<int>()>;

    void setup_output(load_floatn r, load_int itype, load_floatn v, load_float p, load_float rho) {
        load_r = r;
        load_itype = itype;
        load_v = v;
        load_p = p;
        load_rho = rho;
    }
    void start_step(double t) { starts.push_back(t); }
    void update_step(double t, size_t step) {
        times.push_back(t);
        last_step = step;
        if (probe_every && step % probe_every == 0) {
            probe_r = load_r();
            probe_v = load_v();
            probe_rho = load_rho();
            probe_p = load_p();
            probe_itype = load_itype();
            probe_step = step;
        }
    }
    void finish_step() { ++finished; }
    void crash_dump(shared_floatn r, shared_darray<int> itype, shared_floatn v, shared_darray<double> rho,
                    shared_darray<double> p, size_t step, double time) {
        crash_step = step;
        crash_time = time;
        crashed = r && itype && v && rho && p;
    }
    bool crashed = false;
    size_t crash_step = 0;
    double crash_time = 0;

    size_t probe_every = 0;
    std::vector<double> starts, times;
    size_t last_step = 0, finished = 0, probe_step = 0;
    shared_floatn probe_r, probe_v;
    shared_darray<double> probe_rho, probe_p;
    shared_darray<int> probe_itype;
    load_floatn load_r, load_v;
    load_float load_p, load_rho;
    load_int load_itype;
};

}  // namespace shapes
