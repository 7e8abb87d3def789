// Umbrella header of the C++ drop-in API (host code; link libpadenti_hip.so).
#pragma once
#include "image.hpp"
#include "image_loader.hpp"
#include "image_sampler.hpp"
#include "training_set_image.hpp"
#include "training_set.hpp"
#include "tree.hpp"
#include "tree_trainer.hpp"
#include "classifier.hpp"
#include "hip_tree_trainer.hpp"
#include "hip_classifier.hpp"
