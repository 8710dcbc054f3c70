// nbody_addon.cc -- Node-API binding of libnbody_sim.so (include/nbody_sim.h).
//
// The reference's host is TypeScript (src/simulation.ts, src/config.ts);
// this addon lets the JavaScript `Simulation` class in simulation.js call
// the C ABI directly.  Errors from the library become thrown JS Errors with
// nbody_last_error()'s message (the reference throws for an unknown scenario,
// simulation.ts:364).  Typed arrays are passed by pointer, never copied twice.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/nbody_sim.h"

namespace {

#define NAPI_CALL(env, call)                                        \
    do {                                                            \
        if ((call) != napi_ok) {                                    \
            napi_throw_error((env), nullptr, "N-API call failed: " #call); \
            return nullptr;                                         \
        }                                                           \
    } while (0)

napi_value throw_nbody(napi_env env, int rc, nbody_sim *h) {
    std::string msg = "nbody error " + std::to_string(rc) + ": " + nbody_last_error(h);
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
    for (size_t i = argc; i < want; ++i) napi_get_undefined(env, &argv[i]);
    return true;
}

double num(napi_env env, napi_value v, double dflt) {
    napi_valuetype t;
    if (napi_typeof(env, v, &t) != napi_ok || t != napi_number) return dflt;
    double d = dflt;
    napi_get_value_double(env, v, &d);
    return d;
}

double prop(napi_env env, napi_value obj, const char *name, double dflt) {
    napi_valuetype t;
    if (napi_typeof(env, obj, &t) != napi_ok || t != napi_object) return dflt;
    bool has = false;
    if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return dflt;
    napi_value v;
    if (napi_get_named_property(env, obj, name, &v) != napi_ok) return dflt;
    return num(env, v, dflt);
}

void set_num(napi_env env, napi_value obj, const char *name, double v) {
    napi_value x;
    napi_create_double(env, v, &x);
    napi_set_named_property(env, obj, name, x);
}

// Config object (config.ts field names) -> nbody_config.
nbody_config to_config(napi_env env, napi_value obj) {
    nbody_config c;
    nbody_default_config(&c);
    c.grav_constant = (float)prop(env, obj, "gravConstant", c.grav_constant);
    c.bh_theta = (float)prop(env, obj, "bhTheta", c.bh_theta);
    c.substeps = (int32_t)prop(env, obj, "substeps", c.substeps);
    c.delta_time = (float)prop(env, obj, "deltaTime", c.delta_time);
    c.epsilon_multiplier = (float)prop(env, obj, "epsilonMultiplier", c.epsilon_multiplier);
    c.epsilon = (float)prop(env, obj, "epsilon", c.epsilon);
    c.softening_mode = (int32_t)prop(env, obj, "softeningMode", c.softening_mode);
    c.morton_mode = (int32_t)prop(env, obj, "mortonMode", c.morton_mode);
    c.math_mode = (int32_t)prop(env, obj, "mathMode", c.math_mode);
    c.offset = (uint32_t)prop(env, obj, "offset", c.offset);
    c.expansion = (int32_t)prop(env, obj, "expansion", c.expansion);
    return c;
}

void finalize_sim(napi_env, void *data, void *) { nbody_destroy(static_cast<nbody_sim *>(data)); }

nbody_sim *unwrap(napi_env env, napi_value v) {
    void *p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok) return nullptr;
    return static_cast<nbody_sim *>(p);
}

// Float32Array data pointer and element count.
float *f32(napi_env env, napi_value v, size_t *len) {
    napi_typedarray_type type;
    size_t length = 0, offset = 0;
    void *data = nullptr;
    napi_value ab;
    if (napi_get_typedarray_info(env, v, &type, &length, &data, &ab, &offset) != napi_ok ||
        type != napi_float32_array)
        return nullptr;
    *len = length;
    return static_cast<float *>(data);
}

napi_value new_f32(napi_env env, size_t count, float **data) {
    napi_value ab, arr;
    void *raw = nullptr;
    if (napi_create_arraybuffer(env, count * sizeof(float), &raw, &ab) != napi_ok) return nullptr;
    if (napi_create_typedarray(env, napi_float32_array, count, ab, 0, &arr) != napi_ok) return nullptr;
    *data = static_cast<float *>(raw);
    return arr;
}

napi_value DefaultConfig(napi_env env, napi_callback_info) {
    nbody_config c;
    nbody_default_config(&c);
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    set_num(env, o, "gravConstant", c.grav_constant);
    set_num(env, o, "bhTheta", c.bh_theta);
    set_num(env, o, "substeps", c.substeps);
    set_num(env, o, "deltaTime", 0.1 * 1.0 / (60.0 * 1));  // config.ts:13, as a JS double
    set_num(env, o, "epsilonMultiplier", c.epsilon_multiplier);
    return o;
}

napi_value Create(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return nullptr;
    nbody_config c = to_config(env, argv[0]);
    int err = 0;
    nbody_sim *h = nbody_create(&c, (int64_t)num(env, argv[1], 0), (int32_t)num(env, argv[2], 0), &err);
    if (!h) return throw_nbody(env, err, nullptr);
    napi_value ext;
    NAPI_CALL(env, napi_create_external(env, h, finalize_sim, nullptr, &ext));
    return ext;
}

napi_value EnsembleCreate(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return nullptr;
    nbody_config c = to_config(env, argv[0]);
    uint32_t len = 0;
    if (napi_get_array_length(env, argv[1], &len) != napi_ok || len == 0) {
        napi_throw_error(env, nullptr, "nPerSim must be a non-empty array");
        return nullptr;
    }
    int64_t sizes[256];
    if (len > 256) {
        napi_throw_error(env, nullptr, "at most 256 simulations");
        return nullptr;
    }
    for (uint32_t i = 0; i < len; ++i) {
        napi_value e;
        NAPI_CALL(env, napi_get_element(env, argv[1], i, &e));
        sizes[i] = (int64_t)num(env, e, 0);
    }
    int err = 0;
    nbody_sim *h = nbody_ensemble_create(&c, (int32_t)len, sizes, (int32_t)num(env, argv[2], 0), &err);
    if (!h) return throw_nbody(env, err, nullptr);
    napi_value ext;
    NAPI_CALL(env, napi_create_external(env, h, finalize_sim, nullptr, &ext));
    return ext;
}

napi_value SetConfig(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    nbody_config c = to_config(env, argv[1]);
    int rc = nbody_set_config(h, &c);
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

napi_value SetBodies(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    size_t nm = 0, np = 0, nv = 0;
    float *m = f32(env, argv[1], &nm), *p = f32(env, argv[2], &np), *v = f32(env, argv[3], &nv);
    if (!m || !p || !v || np != 2 * nm || nv != 2 * nm) {
        napi_throw_type_error(env, nullptr,
                              "setBodies(mass: Float32Array(n), pos: Float32Array(2n), vel: Float32Array(2n))");
        return nullptr;
    }
    int rc = nbody_set_bodies(h, (int64_t)nm, m, p, v);
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

napi_value GetBodies(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    const int64_t n = nbody_get_num_bodies(h);
    if (n < 0) return throw_nbody(env, NBODY_ERR_INVALID_ARG, h);
    float *m, *p, *v;
    napi_value am = new_f32(env, (size_t)n, &m), ap = new_f32(env, 2 * (size_t)n, &p),
               av = new_f32(env, 2 * (size_t)n, &v);
    if (!am || !ap || !av) return nullptr;
    int rc = nbody_get_bodies(h, m, p, v);
    if (rc) return throw_nbody(env, rc, h);
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    napi_set_named_property(env, o, "mass", am);
    napi_set_named_property(env, o, "pos", ap);
    napi_set_named_property(env, o, "vel", av);
    return o;
}

napi_value Step(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    int rc = nbody_step(h, (int32_t)num(env, argv[1], 1));
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

napi_value Synchronize(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    int rc = nbody_synchronize(h);
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

// useGraph(handle, stepsPerGraph): 0 = launch kernels one by one.
napi_value UseGraph(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    int rc = nbody_use_graph(h, (int32_t)num(env, argv[1], 1));
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

// prepareGraph(handle): capture the step graph now (nbody_prepare_graph).
napi_value PrepareGraph(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    int rc = nbody_prepare_graph(h);
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

// updateRadiusMultiplierBuffer (simulation.ts:382-384).
napi_value SetRadius(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    size_t nr = 0;
    float *r = f32(env, argv[1], &nr);
    if (!r || (int64_t)nr != nbody_get_num_bodies(h)) {
        napi_throw_type_error(env, nullptr, "setRadiusMultiplier(r: Float32Array(n))");
        return nullptr;
    }
    int rc = nbody_set_radius_multiplier(h, r);
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

napi_value GetRadius(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    const int64_t n = nbody_get_num_bodies(h);
    if (n < 0) return throw_nbody(env, NBODY_ERR_INVALID_ARG, h);
    float *r;
    napi_value ar = new_f32(env, (size_t)n, &r);
    if (!ar) return nullptr;
    int rc = nbody_get_radius_multiplier(h, r);
    if (rc) return throw_nbody(env, rc, h);
    return ar;
}

// setScenario's radiusMultiplierData (simulation.ts:240-241, 321, 330, 339).
napi_value GenerateRadius(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    const int64_t n = (int64_t)num(env, argv[1], 0);
    if (n < 0) {
        napi_throw_range_error(env, nullptr, "n must be >= 0");
        return nullptr;
    }
    float *r;
    napi_value ar = new_f32(env, (size_t)n, &r);
    if (!ar) return nullptr;
    if (nbody_generate_radius((int32_t)num(env, argv[0], -1), n, r)) {
        napi_throw_error(env, nullptr, "Unknown scenario");
        return nullptr;
    }
    return ar;
}

std::string str(napi_env env, napi_value v) {
    size_t len = 0;
    napi_get_value_string_utf8(env, v, nullptr, 0, &len);
    std::string out(len, '\0');
    napi_get_value_string_utf8(env, v, &out[0], len + 1, &len);
    return out;
}

// render(handle, {centerX, centerY, halfX, halfY, width, height}) ->
// {rgba: Uint8Array(4wh), density: Uint16Array(wh)} (renderer.ts:293-329).
napi_value Render(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    nbody_camera c;
    c.center_x = (float)prop(env, argv[1], "centerX", 0.0);
    c.center_y = (float)prop(env, argv[1], "centerY", 0.0);
    c.half_x = (float)prop(env, argv[1], "halfX", 10.0);
    c.half_y = (float)prop(env, argv[1], "halfY", 10.0);
    c.width = (int32_t)prop(env, argv[1], "width", 0);
    c.height = (int32_t)prop(env, argv[1], "height", 0);
    if (c.width < 1 || c.height < 1 || c.width > 16384 || c.height > 16384) {
        napi_throw_range_error(env, nullptr, "width and height must be 1..16384");
        return nullptr;
    }
    const size_t px = (size_t)c.width * (size_t)c.height;
    napi_value ab8, ab16, rgba, dens;
    void *p8 = nullptr, *p16 = nullptr;
    NAPI_CALL(env, napi_create_arraybuffer(env, 4 * px, &p8, &ab8));
    NAPI_CALL(env, napi_create_arraybuffer(env, 2 * px, &p16, &ab16));
    NAPI_CALL(env, napi_create_typedarray(env, napi_uint8_array, 4 * px, ab8, 0, &rgba));
    NAPI_CALL(env, napi_create_typedarray(env, napi_uint16_array, px, ab16, 0, &dens));
    int rc = nbody_render(h, &c, static_cast<uint16_t *>(p16), static_cast<uint8_t *>(p8));
    if (rc) return throw_nbody(env, rc, h);
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    napi_set_named_property(env, o, "rgba", rgba);
    napi_set_named_property(env, o, "density", dens);
    return o;
}

// writePng(path, width, height, rgba: Uint8Array)
napi_value WritePng(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return nullptr;
    const int32_t w = (int32_t)num(env, argv[1], 0), hgt = (int32_t)num(env, argv[2], 0);
    napi_typedarray_type type;
    size_t length = 0, offset = 0;
    void *data = nullptr;
    napi_value ab;
    if (napi_get_typedarray_info(env, argv[3], &type, &length, &data, &ab, &offset) != napi_ok ||
        type != napi_uint8_array || w < 1 || hgt < 1 || length != 4 * (size_t)w * (size_t)hgt) {
        napi_throw_type_error(env, nullptr, "writePng(path, width, height, rgba: Uint8Array(4wh))");
        return nullptr;
    }
    int rc = nbody_write_png(str(env, argv[0]).c_str(), w, hgt, static_cast<const uint8_t *>(data));
    if (rc) return throw_nbody(env, rc, nullptr);
    return nullptr;
}

napi_value SaveSnapshot(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    int rc = nbody_save_snapshot(h, str(env, argv[1]).c_str());
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

napi_value LoadSnapshot(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    nbody_sim *h = unwrap(env, argv[0]);
    int rc = nbody_load_snapshot(h, str(env, argv[1]).c_str());
    if (rc) return throw_nbody(env, rc, h);
    return nullptr;
}

napi_value GetNumBodies(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    napi_value r;
    NAPI_CALL(env, napi_create_double(env, (double)nbody_get_num_bodies(unwrap(env, argv[0])), &r));
    return r;
}

napi_value Generate(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return nullptr;
    const int32_t scen = (int32_t)num(env, argv[0], -1);
    const int64_t n = (int64_t)num(env, argv[1], 0);
    const uint64_t seed = (uint64_t)num(env, argv[2], 0);
    nbody_ic_params p{(float)prop(env, argv[3], "gravConstant", 1.0), (float)prop(env, argv[3], "scale", 0.0),
                      (float)prop(env, argv[3], "dispersion", 0.0), (float)prop(env, argv[3], "massSigma", 0.0)};
    if (n < 0) {
        napi_throw_range_error(env, nullptr, "n must be >= 0");
        return nullptr;
    }
    float *m, *ps, *v;
    napi_value am = new_f32(env, (size_t)n, &m), ap = new_f32(env, 2 * (size_t)n, &ps),
               av = new_f32(env, 2 * (size_t)n, &v);
    if (!am || !ap || !av) return nullptr;
    int rc = nbody_generate(scen, n, seed, &p, m, ps, v);
    if (rc) {
        napi_throw_error(env, nullptr, "Unknown scenario");
        return nullptr;
    }
    napi_value o;
    NAPI_CALL(env, napi_create_object(env, &o));
    napi_set_named_property(env, o, "mass", am);
    napi_set_named_property(env, o, "pos", ap);
    napi_set_named_property(env, o, "vel", av);
    return o;
}

napi_value Version(napi_env env, napi_callback_info) {
    napi_value s;
    NAPI_CALL(env, napi_create_string_utf8(env, nbody_version(), NAPI_AUTO_LENGTH, &s));
    return s;
}

napi_value Init(napi_env env, napi_value exports) {
    const napi_property_descriptor props[] = {
        {"defaultConfig", nullptr, DefaultConfig, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"create", nullptr, Create, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"ensembleCreate", nullptr, EnsembleCreate, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"setConfig", nullptr, SetConfig, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"setBodies", nullptr, SetBodies, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"getBodies", nullptr, GetBodies, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"step", nullptr, Step, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"synchronize", nullptr, Synchronize, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"useGraph", nullptr, UseGraph, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"prepareGraph", nullptr, PrepareGraph, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"setRadius", nullptr, SetRadius, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"getRadius", nullptr, GetRadius, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"generateRadius", nullptr, GenerateRadius, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"saveSnapshot", nullptr, SaveSnapshot, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"render", nullptr, Render, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"writePng", nullptr, WritePng, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"loadSnapshot", nullptr, LoadSnapshot, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"getNumBodies", nullptr, GetNumBodies, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"generate", nullptr, Generate, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"version", nullptr, Version, nullptr, nullptr, nullptr, napi_default, nullptr},
    };
    napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
