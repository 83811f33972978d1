// CPython extension `cubnm_amd.core`: the drop-in for the reference's
// `cubnm.core` (src/ext/core.cpp:495-626).  Same three functions with the same
// positional arguments, returned dict and in-place parameter write-back;
// everything below the argument parsing goes through the C ABI
// (include/cubnm_amd.h).
//
// Deliberate deviations from the reference boundary:
//  * the six flags are parsed into int storage (the reference parses "i"
//    into C++ bool objects, core.cpp:198-216);
//  * every failure raises (RuntimeError / TypeError / ValueError) instead of
//    printing and returning None (core.cpp:226-227, 244-252, 271-274, 297-304);
//  * use_cpu=True raises: the CPU engine is not part of this library.
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cubnm_amd.h"

extern "C" const char* cubnm_gof_last_error(void);

namespace {

struct Ref {  // owning PyObject* holder
    PyObject* p = nullptr;
    Ref() = default;
    explicit Ref(PyObject* o) : p(o) {}
    ~Ref() { Py_XDECREF(p); }
    Ref(const Ref&) = delete;
    Ref& operator=(const Ref&) = delete;
    PyArrayObject* arr() const { return reinterpret_cast<PyArrayObject*>(p); }
    PyObject* release() { PyObject* o = p; p = nullptr; return o; }
};

PyObject* raise_core() {
    PyErr_SetString(PyExc_RuntimeError, cubnm_last_error());
    return nullptr;
}

template <class T>
T* data(const Ref& r) { return r.p ? static_cast<T*>(PyArray_DATA(r.arr())) : nullptr; }

PyObject* py_init(PyObject*, PyObject*) {
    cubnm_init();
    Py_RETURN_NONE;
}

PyObject* py_set_const(PyObject*, PyObject* args) {
    const char* key;
    double value;
    if (!PyArg_ParseTuple(args, "sd", &key, &value)) return nullptr;
    cubnm_set_const(key, value);
    Py_RETURN_NONE;
}

PyObject* py_run_simulations(PyObject*, PyObject* args) {
    const char* model_name;
    PyObject *o_SC, *o_SCi, *o_dist, *o_gp, *o_rp, *o_v, *config;
    int ext_out, states_ts, noise_out, do_delay, force_reinit, use_cpu;
    int N_SIMS, nodes, time_steps, BOLD_TR, states_sampling, sim_seed;
    double dt, bw_dt;
    if (!PyArg_ParseTuple(args, "sO!O!O!O!O!O!Oiiiiiiiiiiiidd", &model_name,
                          &PyArray_Type, &o_SC, &PyArray_Type, &o_SCi, &PyArray_Type, &o_dist,
                          &PyArray_Type, &o_gp, &PyArray_Type, &o_rp, &PyArray_Type, &o_v, &config,
                          &ext_out, &states_ts, &noise_out, &do_delay, &force_reinit, &use_cpu,
                          &N_SIMS, &nodes, &time_steps, &BOLD_TR, &states_sampling, &sim_seed, &dt, &bw_dt))
        return nullptr;
    if (!PyDict_Check(config)) {
        PyErr_SetString(PyExc_TypeError, "model_config must be a dictionary.");
        return nullptr;
    }
    std::vector<std::string> keys, values;
    PyObject *k, *v;
    Py_ssize_t pos = 0;
    while (PyDict_Next(config, &pos, &k, &v)) {
        if (!PyUnicode_Check(k) || !PyUnicode_Check(v)) {
            PyErr_SetString(PyExc_TypeError, "Dictionary keys and values must be strings.");
            return nullptr;
        }
        keys.emplace_back(PyUnicode_AsUTF8(k));
        values.emplace_back(PyUnicode_AsUTF8(v));
    }
    // IN_ARRAY returns the caller's array itself when it is already aligned,
    // C-contiguous float64, so the parameter write-back lands in place
    // (core.cpp:230-233, 437-450).
    Ref SC(PyArray_FROM_OTF(o_SC, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    Ref SCi(PyArray_FROM_OTF(o_SCi, NPY_INT32, NPY_ARRAY_IN_ARRAY));
    Ref dist(PyArray_FROM_OTF(o_dist, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    Ref gp(PyArray_FROM_OTF(o_gp, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    Ref rp(PyArray_FROM_OTF(o_rp, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    Ref vl(PyArray_FROM_OTF(o_v, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    if (!SC.p || !SCi.p || !dist.p || !gp.p || !rp.p || !vl.p) return nullptr;
    if (N_SIMS <= 0 || nodes <= 0) {
        PyErr_SetString(PyExc_ValueError, "N_SIMS and nodes must be positive");
        return nullptr;
    }
    if (PyArray_SIZE(SCi.arr()) < N_SIMS || PyArray_SIZE(vl.arr()) < N_SIMS) {
        PyErr_SetString(PyExc_ValueError, "SC_indices and v_list must have N_SIMS entries");
        return nullptr;
    }
    const int* sci = data<int>(SCi);
    int n_scs = 0;
    for (int i = 0; i < N_SIMS; i++) n_scs = std::max(n_scs, sci[i] + 1);
    if (PyArray_SIZE(SC.arr()) < (npy_intp)n_scs * nodes * nodes) {
        PyErr_SetString(PyExc_ValueError, "SC has fewer matrices than max(SC_indices)+1");
        return nullptr;
    }
    if (do_delay && PyArray_SIZE(dist.arr()) < (npy_intp)nodes * nodes) {
        PyErr_SetString(PyExc_ValueError, "SC_dist must have nodes*nodes entries");
        return nullptr;
    }

    std::vector<const char*> kp, vp;
    for (size_t i = 0; i < keys.size(); i++) { kp.push_back(keys[i].c_str()); vp.push_back(values[i].c_str()); }
    cubnm_dims d{};
    if (cubnm_prepare(model_name, (int)kp.size(), kp.data(), vp.data(), ext_out, states_ts, noise_out, do_delay,
                      force_reinit, use_cpu, N_SIMS, nodes, n_scs, time_steps, BOLD_TR, states_sampling, sim_seed,
                      dt, bw_dt, &d) != 0)
        return raise_core();
    if (PyArray_SIZE(gp.arr()) < (npy_intp)d.n_global_params * N_SIMS ||
        PyArray_SIZE(rp.arr()) < (npy_intp)d.n_regional_params * N_SIMS * nodes) {
        PyErr_SetString(PyExc_ValueError, "global/regional parameter arrays are smaller than the model needs");
        return nullptr;
    }

    npy_intp bold_dims[2] = {N_SIMS, (npy_intp)d.bold_len * nodes};
    npy_intp fc_dims[2] = {N_SIMS, d.n_pairs};
    npy_intp fcd_dims[2] = {N_SIMS, d.n_window_pairs};
    npy_intp st_dims[3] = {d.n_state_vars, N_SIMS, (npy_intp)nodes * (states_ts ? d.states_len : 1)};
    npy_intp gb_dims[2] = {d.n_global_out_bool, N_SIMS};
    npy_intp gi_dims[2] = {d.n_global_out_int, N_SIMS};
    npy_intp noise_dims[1] = {(npy_intp)d.noise_size};
    npy_intp shn_dims[2] = {d.noise_repeats, nodes};
    npy_intp sht_dims[2] = {d.noise_repeats, d.noise_bw_it};
    Ref bold(PyArray_SimpleNew(2, bold_dims, NPY_DOUBLE));
    Ref fc, fcd, st, gb, gi, noise, shn, sht;
    if (d.do_fc) fc.p = PyArray_SimpleNew(2, fc_dims, NPY_DOUBLE);
    if (d.do_fcd) fcd.p = PyArray_SimpleNew(2, fcd_dims, NPY_DOUBLE);
    if (ext_out) {
        st.p = PyArray_SimpleNew(3, st_dims, NPY_DOUBLE);
        gb.p = PyArray_ZEROS(2, gb_dims, NPY_BOOL, 0);
        gi.p = PyArray_ZEROS(2, gi_dims, NPY_INT32, 0);
    }
    if (noise_out) {
        noise.p = PyArray_SimpleNew(1, noise_dims, NPY_DOUBLE);
        shn.p = PyArray_SimpleNew(2, shn_dims, NPY_INT32);
        sht.p = PyArray_SimpleNew(2, sht_dims, NPY_INT32);
    }
    if (!bold.p || (d.do_fc && !fc.p) || (d.do_fcd && !fcd.p) || (ext_out && (!st.p || !gb.p || !gi.p)) ||
        (noise_out && (!noise.p || !shn.p || !sht.p)))
        return nullptr;

    cubnm_outputs out{};
    out.sim_bold = data<double>(bold);
    out.sim_fc_trils = data<double>(fc);
    out.sim_fcd_trils = data<double>(fcd);
    out.sim_states = data<double>(st);
    out.global_bools = data<unsigned char>(gb);
    out.global_ints = data<int>(gi);
    out.noise = data<double>(noise);
    out.shuffled_nodes = data<int>(shn);
    out.shuffled_ts = data<int>(sht);
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = cubnm_run(data<double>(SC), sci, data<double>(dist), data<double>(gp), data<double>(rp), data<double>(vl), &out);
    Py_END_ALLOW_THREADS
    if (rc != 0) return raise_core();

    Ref dict(PyDict_New());
    Ref it(PyFloat_FromDouble(d.init_time)), rt(PyFloat_FromDouble(out.run_time));
    PyDict_SetItemString(dict.p, "init_time", it.p);
    PyDict_SetItemString(dict.p, "run_time", rt.p);
    PyDict_SetItemString(dict.p, "sim_bold", bold.p);
    if (d.do_fc) PyDict_SetItemString(dict.p, "sim_fc_trils", fc.p);
    if (d.do_fcd) PyDict_SetItemString(dict.p, "sim_fcd_trils", fcd.p);
    if (ext_out) {
        PyDict_SetItemString(dict.p, "_sim_states", st.p);
        PyDict_SetItemString(dict.p, "_global_bools", gb.p);
        PyDict_SetItemString(dict.p, "_global_ints", gi.p);
    }
    if (noise_out) {
        PyDict_SetItemString(dict.p, "_noise", noise.p);
        PyDict_SetItemString(dict.p, "_shuffled_nodes", shn.p);
        PyDict_SetItemString(dict.p, "_shuffled_ts", sht.p);
    }
    return dict.release();
}

// gof_fc_corr(sim_fc_trils (N, P), emp_fc_tril (P,)) -> (N,) Pearson r
PyObject* py_gof_fc_corr(PyObject*, PyObject* args) {
    PyObject *o_sim, *o_emp;
    if (!PyArg_ParseTuple(args, "OO", &o_sim, &o_emp)) return nullptr;
    Ref sim(PyArray_FROM_OTF(o_sim, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    Ref emp(PyArray_FROM_OTF(o_emp, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    if (!sim.p || !emp.p) return nullptr;
    if (PyArray_NDIM(sim.arr()) != 2 || PyArray_DIM(sim.arr(), 1) != PyArray_SIZE(emp.arr())) {
        PyErr_SetString(PyExc_ValueError, "sim_fc must be (N, n_pairs) and emp_fc (n_pairs,)");
        return nullptr;
    }
    npy_intp n = PyArray_DIM(sim.arr(), 0);
    Ref out(PyArray_SimpleNew(1, &n, NPY_DOUBLE));
    if (cubnm_gof_fc_corr(data<double>(sim), (int)n, (int)PyArray_DIM(sim.arr(), 1), data<double>(emp),
                          data<double>(out)) != 0) {
        PyErr_SetString(PyExc_RuntimeError, cubnm_gof_last_error());
        return nullptr;
    }
    return out.release();
}

// gof_fcd_ks(sim_fcd_trils (N, M1), emp_fcd_tril (M2,)) -> (N,) KS statistic
PyObject* py_gof_fcd_ks(PyObject*, PyObject* args) {
    PyObject *o_sim, *o_emp;
    if (!PyArg_ParseTuple(args, "OO", &o_sim, &o_emp)) return nullptr;
    Ref sim(PyArray_FROM_OTF(o_sim, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    Ref emp(PyArray_FROM_OTF(o_emp, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY));
    if (!sim.p || !emp.p) return nullptr;
    if (PyArray_NDIM(sim.arr()) != 2) {
        PyErr_SetString(PyExc_ValueError, "sim_fcd must be (N, n_window_pairs)");
        return nullptr;
    }
    npy_intp n = PyArray_DIM(sim.arr(), 0);
    Ref out(PyArray_SimpleNew(1, &n, NPY_DOUBLE));
    if (cubnm_gof_fcd_ks(data<double>(sim), (int)n, (int)PyArray_DIM(sim.arr(), 1), data<double>(emp),
                         (int)PyArray_SIZE(emp.arr()), data<double>(out)) != 0) {
        PyErr_SetString(PyExc_RuntimeError, cubnm_gof_last_error());
        return nullptr;
    }
    return out.release();
}

PyObject* py_last_timings(PyObject*, PyObject*) {
    double t[6];
    cubnm_last_timings(t);
    return Py_BuildValue("{s:d,s:d,s:d,s:d,s:d,s:d,s:i,s:i}", "prep_ms", t[0], "h2d_ms", t[1], "integrate_ms", t[2],
                         "fc_fcd_ms", t[3], "d2h_ms", t[4], "integrate_launches", t[5], "integrator",
                         cubnm_last_integrator(), "fc_path", cubnm_last_fc_path());
}

PyObject* py_set_tuning(PyObject*, PyObject* args) {
    const char* key;
    int v;
    if (!PyArg_ParseTuple(args, "si", &key, &v)) return nullptr;
    if (cubnm_set_tuning(key, v) != 0) {
        PyErr_SetString(PyExc_ValueError, cubnm_last_error());
        return nullptr;
    }
    Py_RETURN_NONE;
}

PyObject* py_force_integrator(PyObject*, PyObject* args) {
    int f;
    if (!PyArg_ParseTuple(args, "i", &f)) return nullptr;
    cubnm_force_integrator(f);
    Py_RETURN_NONE;
}

PyMethodDef methods[] = {
    {"run_simulations", py_run_simulations, METH_VARARGS,
     "run_simulations(model_name, SC, SC_indices, SC_dist, global_params, regional_params, v_list,\n"
     "                model_config, ext_out, states_ts, noise_out, do_delay, force_reinit, use_cpu,\n"
     "                N_SIMS, nodes, time_steps, BOLD_TR, states_sampling, sim_seed, dt, bw_dt) -> dict\n\n"
     "Run a group of simulations on the GPU; see include/cubnm_amd.h for layouts."},
    {"init", py_init, METH_NOARGS, "init()\nInitialise the session constants."},
    {"set_const", py_set_const, METH_VARARGS, "set_const(key, value)\nSet BW k1, k2, k3 or V_0."},
    {"gof_fc_corr", py_gof_fc_corr, METH_VARARGS, "gof_fc_corr(sim_fc_trils, emp_fc_tril) -> Pearson r per simulation"},
    {"gof_fcd_ks", py_gof_fcd_ks, METH_VARARGS, "gof_fcd_ks(sim_fcd_trils, emp_fcd_tril) -> KS statistic per simulation"},
    {"last_timings", py_last_timings, METH_NOARGS, "last_timings() -> dict of phase timings of the last run (ms)"},
    {"force_integrator", py_force_integrator, METH_VARARGS, "force_integrator(k): -1 auto, 0 general, 1 batched"},
    {"set_tuning", py_set_tuning, METH_VARARGS,
     "set_tuning(key, value): 'batched_tpw' 0/1/2; 'fc_mode' 0 auto, 1 FP64 Gram, 2 split-BF16 Gram"},
    {nullptr, nullptr, 0, nullptr}};

PyModuleDef module_def = {PyModuleDef_HEAD_INIT, "core", "MI355X simulation core (drop-in for cubnm.core)", -1, methods};

}  // namespace

PyMODINIT_FUNC PyInit_core(void) {
    import_array();
    return PyModule_Create(&module_def);
}
