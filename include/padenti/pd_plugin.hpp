// Custom per-pixel feature hooks (the HIP replacement of the reference's
// feature.cl plugin: a directory passed to CLTreeTrainer/CLClassifier and
// #included into the kernels, cl_tree_trainer_impl.hpp:119-192,
// test/feature.cl:29-34).
//
// Write a functor with the interface documented in pd_features.hpp, then in
// a .hip file compiled for gfx950:
//
//   #include <padenti/pd_plugin.hpp>
//   struct MyFeature { ... };
//   PD_DEFINE_FEATURE(my_feature, MyFeature)
//
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared -ffp-contract=off \
//         -I<repo>/include my_feature.hip -o libmy_feature.so
//
// The library exports
//   int pd_trainer_create_my_feature(pd_trainer **out, int device, uint32_t n_classes);
//   int pd_forest_create_my_feature(pd_forest **out, int device, uint32_t n_classes);
//   const char *pd_last_error_my_feature(void);
// and every other entry point of padenti_hip.h (set_data, train, add_tree,
// predict, destroy, ...) works on the returned handles.
#pragma once
#include "pd_engine.hpp"

#define PD_DEFINE_FEATURE(name, F)                                                                      \
  extern "C" PD_API int pd_trainer_create_##name(pd_trainer **out, int device, uint32_t n_classes)            \
  {                                                                                                     \
    return pd::engine::guarded([&] {                                                                    \
      pd::engine::require(out != nullptr, "null output handle");                                      \
      pd::engine::require(n_classes >= 1 && n_classes <= 64, "n_classes must be in [1, 64]");         \
      std::unique_ptr<pd_trainer> t(new pd_trainer);                                                    \
      t->impl.reset(new pd::engine::Trainer<F>(device, n_classes));                                    \
      *out = t.release();                                                                               \
    });                                                                                                 \
  }                                                                                                     \
  extern "C" PD_API int pd_forest_create_##name(pd_forest **out, int device, uint32_t n_classes)              \
  {                                                                                                     \
    return pd::engine::guarded([&] {                                                                    \
      pd::engine::require(out != nullptr, "null output handle");                                      \
      pd::engine::require(n_classes >= 1 && n_classes <= 64, "n_classes must be in [1, 64]");         \
      std::unique_ptr<pd_forest> f(new pd_forest);                                                      \
      f->impl.reset(new pd::engine::Forest<F>(device, n_classes));                                     \
      *out = f.release();                                                                               \
    });                                                                                                 \
  }                                                                                                     \
  extern "C" PD_API const char *pd_last_error_##name(void) { return pd::engine::g_err.c_str(); }
