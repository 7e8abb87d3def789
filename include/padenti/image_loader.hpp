// Image loaders: the ImageLoader interface (src/padenti/image_loader.hpp:
// 27-55) and PNG stand-ins for the OpenCV loaders
// (src/padenti/cv_image_loader_impl.hpp:36-133):
//   PNGImageLoader<type, nChannels>  CVImageLoader: the file must decode to
//                                    exactly this element type and channel
//                                    count (OpenCV's B, G, R(, A) order)
//   RGBLabelsLoader                  CVRGBLabelsLoader: an 8-bit RGB(A) PNG
//                                    mapped to labels 1..nClasses (0 when no
//                                    colour matches) on the GPU
// Decoding runs on the host (libpadenti_hip.so's PNG codec).  Errors throw
// std::runtime_error with the reference's messages where it has one.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../padenti_hip.h"
#include "image.hpp"

template <typename type, unsigned int nChannels>
class ImageLoader
{
public:
  virtual ~ImageLoader() {}
  // data must hold width * height * nChannels values
  virtual void load(const std::string &imagePath, type *data, unsigned int *width, unsigned int *height) = 0;
  virtual Image<type, nChannels> load(const std::string &imagePath) = 0;
};

namespace pd_cxx {
inline void io_check(int rc)
{
  if (rc < 0) throw std::runtime_error(std::string("padenti_hip: ") + pd_last_error());
}
// bit depth a PNG decodes to for an element type (0: never, as OpenCV's
// depth check fails for every other type)
template <typename T> struct png_depth { static const unsigned int value = 0; };
template <> struct png_depth<unsigned char> { static const unsigned int value = 8; };
template <> struct png_depth<unsigned short> { static const unsigned int value = 16; };
}  // namespace pd_cxx

template <typename type, unsigned int nChannels>
class PNGImageLoader : public ImageLoader<type, nChannels>
{
  static pd_png_info header(const std::string &path)
  {
    pd_png_info info;
    pd_cxx::io_check(pd_png_read(path.c_str(), &info, NULL, 0));
    if (info.bit_depth != pd_cxx::png_depth<type>::value || info.channels != nChannels)
      throw std::runtime_error("Image type different from expected one");
    return info;
  }

public:
  void load(const std::string &imagePath, type *data, unsigned int *width, unsigned int *height)
  {
    pd_png_info info = header(imagePath);
    pd_cxx::io_check(pd_png_read(imagePath.c_str(), &info, data,
                                 (uint64_t)info.width * info.height * nChannels * sizeof(type)));
    *width = info.width;
    *height = info.height;
  }
  Image<type, nChannels> load(const std::string &imagePath)
  {
    pd_png_info info = header(imagePath);
    Image<type, nChannels> img(info.width, info.height);
    pd_cxx::io_check(pd_png_read(imagePath.c_str(), &info, img.getData(),
                                 (uint64_t)info.width * info.height * nChannels * sizeof(type)));
    return img;
  }
};

class RGBLabelsLoader : public ImageLoader<unsigned char, 1>
{
  std::vector<unsigned char> m_map;  // [nClasses][3] R, G, B
  unsigned int m_nClasses;
  int m_device;

  void decode(const std::string &path, std::vector<unsigned char> &labels, unsigned int *w, unsigned int *h) const
  {
    pd_png_info info;
    pd_cxx::io_check(pd_png_read(path.c_str(), &info, NULL, 0));
    if (info.bit_depth != 8 || (info.channels != 3 && info.channels != 4))
      throw std::runtime_error("Labels image type must be CV_8UC3 or CV_8UC4");
    const size_t n = (size_t)info.width * info.height;
    std::vector<unsigned char> px(n * info.channels);
    pd_cxx::io_check(pd_png_read(path.c_str(), &info, px.data(), px.size()));
    for (size_t i = 0; i < n; i++) std::swap(px[i * info.channels], px[i * info.channels + 2]);  // BGR -> RGB
    labels.resize(n);
    pd_cxx::io_check(pd_rgb_to_labels(m_device, n, info.channels, px.data(), m_map.data(), m_nClasses, labels.data()));
    *w = info.width;
    *h = info.height;
  }

public:
  RGBLabelsLoader(const unsigned char (*rgb2labelMap)[3], unsigned int nClasses, int device = 0)
    : m_map(&rgb2labelMap[0][0], &rgb2labelMap[0][0] + 3 * (size_t)nClasses), m_nClasses(nClasses), m_device(device)
  {
  }
  void load(const std::string &imagePath, unsigned char *data, unsigned int *width, unsigned int *height)
  {
    std::vector<unsigned char> lab;
    decode(imagePath, lab, width, height);
    std::copy(lab.begin(), lab.end(), data);
  }
  Image<unsigned char, 1> load(const std::string &imagePath)
  {
    std::vector<unsigned char> lab;
    unsigned int w, h;
    decode(imagePath, lab, &w, &h);
    return Image<unsigned char, 1>(lab.data(), w, h);
  }
};
