// Tree<FeatType, FeatDim, nClasses>: breadth-first heap in SoA arrays
// (src/padenti/tree.hpp:31-173, tree_impl.hpp:239-273): node i has children
// 2i+1 / 2i+2; leftChild is -2 for a nonexistent node, -1 for a leaf and
// 2i+1 for a split node.  A tree of depth D has 2^D - 1 nodes.
//
// load/save use the reference's XML layout (tree_impl.hpp:103-234, Boost
// property_tree write_xml): <Trees><TreeNumber>1</TreeNumber>
// <MaxDepth>D</MaxDepth><Tree{id}><Node{i}><LeftChild>..</LeftChild>
// <SplitParameters>f0 .. fFD-1 thr</SplitParameters><Histogram>p0 .. pC-1
// </Histogram></Node{i}>..</Tree{id}></Trees>, nonexistent nodes omitted,
// numbers written with the default stream precision (6 significant digits)
// unless `lossless` asks for max_digits10.
#pragma once
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

template <typename FeatType, unsigned int FeatDim>
class TreeNode
{
public:
  int *m_leftChild;
  FeatType *m_feature;
  FeatType *m_threshold;
  float *m_posterior;
};

namespace pd_xml {

// Minimal reader for the element-only XML written by save(): every
// element's text, keyed by its dotted path ("Trees.Tree0.Node3.LeftChild").
inline void parse(const std::string &s, size_t &i, const std::string &path, std::map<std::string, std::string> &out)
{
  while (i < s.size()) {
    const size_t lt = s.find('<', i);
    if (lt == std::string::npos) return;
    if (s.compare(lt, 2, "<?") == 0 || s.compare(lt, 4, "<!--") == 0) {
      const size_t e = s.find('>', lt);
      if (e == std::string::npos) throw std::runtime_error("malformed tree XML");
      i = e + 1;
      continue;
    }
    if (s.compare(lt, 2, "</") == 0) {
      const size_t e = s.find('>', lt);
      if (e == std::string::npos) throw std::runtime_error("malformed tree XML");
      i = e + 1;
      return;
    }
    const size_t gt = s.find('>', lt);
    if (gt == std::string::npos) throw std::runtime_error("malformed tree XML");
    std::string name = s.substr(lt + 1, gt - lt - 1);
    const bool empty = !name.empty() && name[name.size() - 1] == '/';
    if (empty) name.erase(name.size() - 1);
    const std::string p = path.empty() ? name : path + "." + name;
    i = gt + 1;
    if (empty) {
      out[p] = "";
      continue;
    }
    const size_t next = s.find('<', i);
    if (next != std::string::npos && s.compare(next, 2, "</") == 0) {
      out[p] = s.substr(i, next - i);
      i = s.find('>', next) + 1;
    } else {
      parse(s, i, p, out);
    }
  }
}

}  // namespace pd_xml

template <typename FeatType, unsigned int FeatDim, unsigned int nClasses>
class Tree
{
private:
  unsigned int m_id;
  unsigned int m_depth;
  std::vector<TreeNode<FeatType, FeatDim> > m_nodes;
  std::vector<int> m_leftChildren;
  std::vector<FeatType> m_features;
  std::vector<FeatType> m_thresholds;
  std::vector<float> m_posteriors;

  void _init()
  {
    const size_t n = ((size_t)1 << m_depth) - 1;
    m_leftChildren.assign(n, -2);
    m_features.assign(n * FeatDim, (FeatType)0);
    m_thresholds.assign(n, (FeatType)0);
    m_posteriors.assign(n * nClasses, 0.0f);
    m_nodes.resize(n);
    for (size_t i = 0; i < n; i++) {
      m_nodes[i].m_leftChild = &m_leftChildren[i];
      m_nodes[i].m_feature = &m_features[i * FeatDim];
      m_nodes[i].m_threshold = &m_thresholds[i];
      m_nodes[i].m_posterior = &m_posteriors[i * nClasses];
    }
  }

  Tree(const Tree &);             // non-copyable (the reference's shallow copy double-frees)
  Tree &operator=(const Tree &);

public:
  Tree() : m_id(0), m_depth(0) {}
  Tree(unsigned int id, unsigned int depth) : m_id(id), m_depth(depth) { _init(); }
  unsigned int getID() const { return m_id; }
  unsigned int getDepth() const { return m_depth; }
  size_t getNNodes() const { return m_leftChildren.size(); }
  const TreeNode<FeatType, FeatDim> &getNode(unsigned int idx) const { return m_nodes[idx]; }
  int *getLeftChildren() const { return const_cast<int *>(m_leftChildren.data()); }
  FeatType *getFeatures() const { return const_cast<FeatType *>(m_features.data()); }
  FeatType *getThresholds() const { return const_cast<FeatType *>(m_thresholds.data()); }
  float *getPosteriors() const { return const_cast<float *>(m_posteriors.data()); }

  void load(const std::string &treePath, int idx = -1)
  {
    std::ifstream f(treePath.c_str());
    if (!f) throw std::runtime_error("cannot open " + treePath);
    std::stringstream ss;
    ss << f.rdbuf();
    std::map<std::string, std::string> kv;
    size_t i = 0;
    const std::string text = ss.str();
    pd_xml::parse(text, i, "", kv);
    if (idx != -1) m_id = (unsigned int)idx;
    if (!kv.count("Trees.MaxDepth")) throw std::runtime_error("tree XML without Trees.MaxDepth");
    const unsigned int depth = (unsigned int)std::stoul(kv["Trees.MaxDepth"]);
    if (!m_depth || m_depth != depth) {
      m_depth = depth;
      _init();
    }
    for (size_t n = 0; n < m_leftChildren.size(); n++) {
      std::ostringstream base;
      base << "Trees.Tree" << m_id << ".Node" << n;
      const std::map<std::string, std::string>::const_iterator lc = kv.find(base.str() + ".LeftChild");
      m_leftChildren[n] = lc == kv.end() ? -2 : std::stoi(lc->second);
      std::fill(&m_features[n * FeatDim], &m_features[n * FeatDim] + FeatDim, (FeatType)0);
      m_thresholds[n] = (FeatType)0;
      std::fill(&m_posteriors[n * nClasses], &m_posteriors[n * nClasses] + nClasses, 0.0f);
      if (m_leftChildren[n] == -2) continue;
      std::istringstream hs(kv[base.str() + ".Histogram"]);
      for (unsigned int c = 0; c < nClasses; c++) hs >> m_posteriors[n * nClasses + c];
      if (m_leftChildren[n] >= 0) {
        std::istringstream ps(kv[base.str() + ".SplitParameters"]);
        for (unsigned int j = 0; j < FeatDim; j++) read_value(ps, m_features[n * FeatDim + j]);
        read_value(ps, m_thresholds[n]);
      }
    }
  }

  void save(const std::string &treePath, int idx = -1, bool lossless = false) const
  {
    std::ostringstream x;
    if (lossless) x.precision(std::numeric_limits<float>::max_digits10);
    x << "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<Trees><TreeNumber>1</TreeNumber><MaxDepth>" << m_depth
      << "</MaxDepth><Tree" << (idx != -1 ? (unsigned int)idx : m_id) << ">";
    for (size_t n = 0; n < m_leftChildren.size(); n++) {
      if (m_leftChildren[n] == -2) continue;
      x << "<Node" << n << "><LeftChild>" << m_leftChildren[n] << "</LeftChild>";
      if (m_leftChildren[n] != -1) {
        x << "<SplitParameters>";
        for (unsigned int j = 0; j < FeatDim; j++) write_value(x, m_features[n * FeatDim + j]) << " ";
        write_value(x, m_thresholds[n]) << "</SplitParameters>";
      }
      x << "<Histogram>";
      for (unsigned int c = 0; c < nClasses; c++) x << m_posteriors[n * nClasses + c] << (c + 1 < nClasses ? " " : "");
      x << "</Histogram></Node" << n << ">";
    }
    x << "</Tree" << (idx != -1 ? (unsigned int)idx : m_id) << "></Trees>";
    std::ofstream f(treePath.c_str());
    if (!f) throw std::runtime_error("cannot write " + treePath);
    f << x.str();
  }

private:
  // short / char feature types stream as numbers, not characters
  static std::ostream &write_value(std::ostream &o, FeatType v)
  {
    if (sizeof(FeatType) < sizeof(int) && std::numeric_limits<FeatType>::is_integer) return o << (int)v;
    return o << v;
  }
  static void read_value(std::istream &i, FeatType &v)
  {
    if (sizeof(FeatType) < sizeof(int) && std::numeric_limits<FeatType>::is_integer) {
      int t = 0;
      i >> t;
      v = (FeatType)t;
    } else {
      i >> v;
    }
  }
};
