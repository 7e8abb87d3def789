// TreeTrainerParameters and the TreeTrainer interface
// (src/padenti/tree_trainer.hpp:35-86).
#pragma once
#include "training_set.hpp"
#include "tree.hpp"

template <typename FeatType, unsigned int FeatDim>
class TreeTrainerParameters
{
public:
  unsigned int nFeatures;          // candidate features per node
  unsigned int nThresholds;        // thresholds per candidate
  bool computeFRange;              // unused (as in the reference)
  FeatType featLowBounds[FeatDim]; // lower bound of each feature entry
  FeatType featUpBounds[FeatDim];  // upper bound of each feature entry
  FeatType thrLowBound;            // lower threshold bound
  FeatType thrUpBound;             // upper threshold bound
  bool randomThrSampling;          // unused (as in the reference)
  unsigned int perLeafSamplesThr;  // a node splits only with more samples than this
};

template <typename ImgType, unsigned int nChannels, typename FeatType, unsigned int FeatDim, unsigned int nClasses>
class TreeTrainer
{
public:
  virtual void train(Tree<FeatType, FeatDim, nClasses> &tree, const TrainingSet<ImgType, nChannels> &trainingSet,
                     const TreeTrainerParameters<FeatType, FeatDim> &params, unsigned int startDepth,
                     unsigned int endDepth) = 0;
  virtual ~TreeTrainer() {}
};
