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
#include <memory>
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

// One run_simulations request: parsed arguments, output arrays, result dict.
struct Call {
    std::string model;
    std::vector<std::string> keys, values;
    std::vector<const char*> kp, vp;
    Ref SC, SCi, dist, gp, rp, vl;
    int ext_out = 0, states_ts = 0, noise_out = 0, do_delay = 0, force_reinit = 0, use_cpu = 0;
    int N_SIMS = 0, nodes = 0, time_steps = 0, BOLD_TR = 0, states_sampling = 0, sim_seed = 0, n_scs = 0;
    double dt = 0, bw_dt = 0;
    cubnm_dims d{};
    Ref bold, fc, fcd, st, gb, gi, noise, shn, sht;
    cubnm_outputs out{};
};

bool parse_call(PyObject* args, Call& c) {
    const char* model_name;
    PyObject *o_SC, *o_SCi, *o_dist, *o_gp, *o_rp, *o_v, *config;
    if (!PyArg_ParseTuple(args, "sO!O!O!O!O!O!Oiiiiiiiiiiiidd", &model_name,
                          &PyArray_Type, &o_SC, &PyArray_Type, &o_SCi, &PyArray_Type, &o_dist,
                          &PyArray_Type, &o_gp, &PyArray_Type, &o_rp, &PyArray_Type, &o_v, &config,
                          &c.ext_out, &c.states_ts, &c.noise_out, &c.do_delay, &c.force_reinit, &c.use_cpu,
                          &c.N_SIMS, &c.nodes, &c.time_steps, &c.BOLD_TR, &c.states_sampling, &c.sim_seed, &c.dt,
                          &c.bw_dt))
        return false;
    c.model = model_name;
    if (!PyDict_Check(config)) {
        PyErr_SetString(PyExc_TypeError, "model_config must be a dictionary.");
        return false;
    }
    PyObject *k, *v;
    Py_ssize_t pos = 0;
    while (PyDict_Next(config, &pos, &k, &v)) {
        if (!PyUnicode_Check(k) || !PyUnicode_Check(v)) {
            PyErr_SetString(PyExc_TypeError, "Dictionary keys and values must be strings.");
            return false;
        }
        c.keys.emplace_back(PyUnicode_AsUTF8(k));
        c.values.emplace_back(PyUnicode_AsUTF8(v));
    }
    for (size_t i = 0; i < c.keys.size(); i++) { c.kp.push_back(c.keys[i].c_str()); c.vp.push_back(c.values[i].c_str()); }
    // IN_ARRAY returns the caller's array itself when it is already aligned,
    // C-contiguous float64, so the parameter write-back lands in place
    // (core.cpp:230-233, 437-450).
    c.SC.p = PyArray_FROM_OTF(o_SC, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    c.SCi.p = PyArray_FROM_OTF(o_SCi, NPY_INT32, NPY_ARRAY_IN_ARRAY);
    c.dist.p = PyArray_FROM_OTF(o_dist, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    c.gp.p = PyArray_FROM_OTF(o_gp, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    c.rp.p = PyArray_FROM_OTF(o_rp, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    c.vl.p = PyArray_FROM_OTF(o_v, NPY_DOUBLE, NPY_ARRAY_IN_ARRAY);
    if (!c.SC.p || !c.SCi.p || !c.dist.p || !c.gp.p || !c.rp.p || !c.vl.p) return false;
    if (c.N_SIMS <= 0 || c.nodes <= 0) {
        PyErr_SetString(PyExc_ValueError, "N_SIMS and nodes must be positive");
        return false;
    }
    if (PyArray_SIZE(c.SCi.arr()) < c.N_SIMS || PyArray_SIZE(c.vl.arr()) < c.N_SIMS) {
        PyErr_SetString(PyExc_ValueError, "SC_indices and v_list must have N_SIMS entries");
        return false;
    }
    const int* sci = data<int>(c.SCi);
    for (int i = 0; i < c.N_SIMS; i++) c.n_scs = std::max(c.n_scs, sci[i] + 1);
    if (PyArray_SIZE(c.SC.arr()) < (npy_intp)c.n_scs * c.nodes * c.nodes) {
        PyErr_SetString(PyExc_ValueError, "SC has fewer matrices than max(SC_indices)+1");
        return false;
    }
    if (c.do_delay && PyArray_SIZE(c.dist.arr()) < (npy_intp)c.nodes * c.nodes) {
        PyErr_SetString(PyExc_ValueError, "SC_dist must have nodes*nodes entries");
        return false;
    }
    return true;
}

// After the dimensions are known: check parameter sizes, allocate outputs.
bool alloc_outputs(Call& c) {
    const cubnm_dims& d = c.d;
    if (PyArray_SIZE(c.gp.arr()) < (npy_intp)d.n_global_params * c.N_SIMS ||
        PyArray_SIZE(c.rp.arr()) < (npy_intp)d.n_regional_params * c.N_SIMS * c.nodes) {
        PyErr_SetString(PyExc_ValueError, "global/regional parameter arrays are smaller than the model needs");
        return false;
    }
    npy_intp bold_dims[2] = {c.N_SIMS, (npy_intp)d.bold_len * c.nodes};
    npy_intp fc_dims[2] = {c.N_SIMS, d.n_pairs};
    npy_intp fcd_dims[2] = {c.N_SIMS, d.n_window_pairs};
    npy_intp st_dims[3] = {d.n_state_vars, c.N_SIMS, (npy_intp)c.nodes * (c.states_ts ? d.states_len : 1)};
    npy_intp gb_dims[2] = {d.n_global_out_bool, c.N_SIMS};
    npy_intp gi_dims[2] = {d.n_global_out_int, c.N_SIMS};
    npy_intp noise_dims[1] = {(npy_intp)d.noise_size};
    npy_intp shn_dims[2] = {d.noise_repeats, c.nodes};
    npy_intp sht_dims[2] = {d.noise_repeats, d.noise_bw_it};
    c.bold.p = PyArray_SimpleNew(2, bold_dims, NPY_DOUBLE);
    if (d.do_fc) c.fc.p = PyArray_SimpleNew(2, fc_dims, NPY_DOUBLE);
    if (d.do_fcd) c.fcd.p = PyArray_SimpleNew(2, fcd_dims, NPY_DOUBLE);
    if (c.ext_out) {
        c.st.p = PyArray_SimpleNew(3, st_dims, NPY_DOUBLE);
        c.gb.p = PyArray_ZEROS(2, gb_dims, NPY_BOOL, 0);
        c.gi.p = PyArray_ZEROS(2, gi_dims, NPY_INT32, 0);
    }
    if (c.noise_out) {
        c.noise.p = PyArray_SimpleNew(1, noise_dims, NPY_DOUBLE);
        c.shn.p = PyArray_SimpleNew(2, shn_dims, NPY_INT32);
        c.sht.p = PyArray_SimpleNew(2, sht_dims, NPY_INT32);
    }
    if (!c.bold.p || (d.do_fc && !c.fc.p) || (d.do_fcd && !c.fcd.p) || (c.ext_out && (!c.st.p || !c.gb.p || !c.gi.p)) ||
        (c.noise_out && (!c.noise.p || !c.shn.p || !c.sht.p)))
        return false;
    c.out.sim_bold = data<double>(c.bold);
    c.out.sim_fc_trils = data<double>(c.fc);
    c.out.sim_fcd_trils = data<double>(c.fcd);
    c.out.sim_states = data<double>(c.st);
    c.out.global_bools = data<unsigned char>(c.gb);
    c.out.global_ints = data<int>(c.gi);
    c.out.noise = data<double>(c.noise);
    c.out.shuffled_nodes = data<int>(c.shn);
    c.out.shuffled_ts = data<int>(c.sht);
    return true;
}

// The reference's returned dict (core.cpp:461-492).
PyObject* result_dict(Call& c) {
    Ref dict(PyDict_New());
    Ref it(PyFloat_FromDouble(c.d.init_time)), rt(PyFloat_FromDouble(c.out.run_time));
    PyDict_SetItemString(dict.p, "init_time", it.p);
    PyDict_SetItemString(dict.p, "run_time", rt.p);
    PyDict_SetItemString(dict.p, "sim_bold", c.bold.p);
    if (c.d.do_fc) PyDict_SetItemString(dict.p, "sim_fc_trils", c.fc.p);
    if (c.d.do_fcd) PyDict_SetItemString(dict.p, "sim_fcd_trils", c.fcd.p);
    if (c.ext_out) {
        PyDict_SetItemString(dict.p, "_sim_states", c.st.p);
        PyDict_SetItemString(dict.p, "_global_bools", c.gb.p);
        PyDict_SetItemString(dict.p, "_global_ints", c.gi.p);
    }
    if (c.noise_out) {
        PyDict_SetItemString(dict.p, "_noise", c.noise.p);
        PyDict_SetItemString(dict.p, "_shuffled_nodes", c.shn.p);
        PyDict_SetItemString(dict.p, "_shuffled_ts", c.sht.p);
    }
    return dict.release();
}

PyObject* py_run_simulations(PyObject*, PyObject* args) {
    Call c;
    if (!parse_call(args, c)) return nullptr;
    if (cubnm_prepare(c.model.c_str(), (int)c.kp.size(), c.kp.data(), c.vp.data(), c.ext_out, c.states_ts,
                      c.noise_out, c.do_delay, c.force_reinit, c.use_cpu, c.N_SIMS, c.nodes, c.n_scs, c.time_steps,
                      c.BOLD_TR, c.states_sampling, c.sim_seed, c.dt, c.bw_dt, &c.d) != 0)
        return raise_core();
    if (!alloc_outputs(c)) return nullptr;
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = cubnm_run(data<double>(c.SC), data<int>(c.SCi), data<double>(c.dist), data<double>(c.gp), data<double>(c.rp),
                   data<double>(c.vl), &c.out);
    Py_END_ALLOW_THREADS
    if (rc != 0) return raise_core();
    return result_dict(c);
}

PyObject* py_run_groups_info(PyObject*, PyObject*) {
    double t[6];
    cubnm_run_groups_info(t);
    return Py_BuildValue("{s:i,s:i,s:i,s:i,s:d,s:i}", "graph", (int)t[0], "ksplit", (int)t[1], "pair", (int)t[2],
                         "tuned", (int)t[3], "trial_ms", t[4], "shapes_timed", (int)t[5]);
}

// run_groups(requests, batched_tpw=None) -> list of dicts: each request is a
// run_simulations argument tuple; all run in one call (cubnm_run_groups).
PyObject* py_run_groups(PyObject*, PyObject* args) {
    PyObject *reqs, *tpws = Py_None;
    if (!PyArg_ParseTuple(args, "O|O", &reqs, &tpws)) return nullptr;
    Ref seq(PySequence_Fast(reqs, "requests must be a sequence of run_simulations argument tuples"));
    if (!seq.p) return nullptr;
    const Py_ssize_t n = PySequence_Fast_GET_SIZE(seq.p);
    if (n == 0) return PyList_New(0);
    std::vector<std::unique_ptr<Call>> calls;
    std::vector<cubnm_group> groups(n);
    for (Py_ssize_t i = 0; i < n; i++) {
        calls.emplace_back(new Call());
        PyObject* t = PySequence_Fast_GET_ITEM(seq.p, i);
        if (!PyTuple_Check(t)) { PyErr_SetString(PyExc_TypeError, "each request must be a tuple"); return nullptr; }
        if (!parse_call(t, *calls[i])) return nullptr;
        Call& c = *calls[i];
        if (c.use_cpu) { PyErr_SetString(PyExc_RuntimeError, "use_cpu is not supported"); return nullptr; }
        cubnm_group& g = groups[i];
        g = cubnm_group{};
        g.model_name = c.model.c_str(); g.n_config = (int)c.kp.size(); g.config_keys = c.kp.data();
        g.config_values = c.vp.data(); g.ext_out = c.ext_out; g.states_ts = c.states_ts; g.noise_out = c.noise_out;
        g.do_delay = c.do_delay; g.force_reinit = c.force_reinit; g.N_SIMS = c.N_SIMS; g.nodes = c.nodes;
        g.N_SCs = c.n_scs; g.time_steps = c.time_steps; g.BOLD_TR = c.BOLD_TR; g.states_sampling = c.states_sampling;
        g.sim_seed = c.sim_seed; g.dt = c.dt; g.bw_dt = c.bw_dt;
        if (tpws != Py_None) {
            PyObject* o = PySequence_GetItem(tpws, i);
            if (!o) return nullptr;
            g.batched_tpw = (int)PyLong_AsLong(o);
            Py_DECREF(o);
            if (PyErr_Occurred()) return nullptr;
        }
        g.SC = data<double>(c.SC); g.SC_indices = data<int>(c.SCi); g.SC_dist = data<double>(c.dist);
        g.global_params = data<double>(c.gp); g.regional_params = data<double>(c.rp); g.v_list = data<double>(c.vl);
    }
    if (cubnm_prepare_groups((int)n, groups.data()) != 0) return raise_core();
    for (Py_ssize_t i = 0; i < n; i++) {
        calls[i]->d = groups[i].dims;
        if (!alloc_outputs(*calls[i])) return nullptr;
        groups[i].out = calls[i]->out;
    }
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = cubnm_run_groups((int)n, groups.data());
    Py_END_ALLOW_THREADS
    if (rc != 0) return raise_core();
    Ref list(PyList_New(n));
    if (!list.p) return nullptr;
    for (Py_ssize_t i = 0; i < n; i++) {
        calls[i]->out.run_time = groups[i].out.run_time;
        PyObject* dct = result_dict(*calls[i]);
        if (!dct) return nullptr;
        PyList_SET_ITEM(list.p, i, dct);
        // integration kernel time of the launch that ran this request (ms)
        double t[6];
        if (cubnm_group_timings((int)i, t) == 0) {
            Ref im(PyFloat_FromDouble(t[2]));
            if (!im.p || PyDict_SetItemString(dct, "integrate_ms", im.p) != 0) return nullptr;
        }
    }
    return list.release();
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
    {"run_groups_info", py_run_groups_info, METH_NOARGS,
     "run_groups_info() -> dict: graph (0 none, 1 captured, 2 replayed), ksplit, pair, tuned (0 rule, 1 trials,\n"
     "2 cached), trial_ms, shapes_timed of the last run_groups call's per-simulation launch"},
    {"run_groups", py_run_groups, METH_VARARGS,
     "run_groups(requests, batched_tpw=None) -> list of dicts\n\n"
     "Run several run_simulations requests (each its own model, node count, duration, config, SCs\n"
     "and parameters) in one call; outputs as run_simulations."},
    {"init", py_init, METH_NOARGS, "init()\nInitialise the session constants."},
    {"set_const", py_set_const, METH_VARARGS, "set_const(key, value)\nSet BW k1, k2, k3 or V_0."},
    {"gof_fc_corr", py_gof_fc_corr, METH_VARARGS, "gof_fc_corr(sim_fc_trils, emp_fc_tril) -> Pearson r per simulation"},
    {"gof_fcd_ks", py_gof_fcd_ks, METH_VARARGS, "gof_fcd_ks(sim_fcd_trils, emp_fcd_tril) -> KS statistic per simulation"},
    {"last_timings", py_last_timings, METH_NOARGS, "last_timings() -> dict of phase timings of the last run (ms)"},
    {"force_integrator", py_force_integrator, METH_VARARGS, "force_integrator(k): -1 auto, 0 general, 1 batched"},
    {"set_tuning", py_set_tuning, METH_VARARGS,
     "set_tuning(key, value): 'batched_tpw' 0/1/2; 'general_ksplit' 0 auto or 1..8; "
     "'fc_mode' 0 auto, 1 FP64 Gram, 2 split-BF16 Gram"},
    {nullptr, nullptr, 0, nullptr}};

PyModuleDef module_def = {PyModuleDef_HEAD_INIT, "core", "MI355X simulation core (drop-in for cubnm.core)", -1, methods};

}  // namespace

PyMODINIT_FUNC PyInit_core(void) {
    import_array();
    return PyModule_Create(&module_def);
}
