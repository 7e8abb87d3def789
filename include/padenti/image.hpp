// Image<type, nChannels>: interleaved pixel buffer (src/padenti/image.hpp:34-111).
#pragma once
#include <algorithm>
#include <cstddef>

template <typename type, unsigned int nChannels>
class Image
{
protected:
  type *m_data;
  unsigned int m_width;
  unsigned int m_height;

public:
  Image() : m_data(NULL), m_width(0), m_height(0) {}
  Image(const type *data, unsigned int width, unsigned int height)
    : m_data(new type[(size_t)width * height * nChannels]), m_width(width), m_height(height)
  {
    std::copy(data, data + (size_t)width * height * nChannels, m_data);
  }
  Image(unsigned int width, unsigned int height)
    : m_data(new type[(size_t)width * height * nChannels]), m_width(width), m_height(height)
  {
  }
  Image(const Image &o) : m_data(new type[(size_t)o.m_width * o.m_height * nChannels]), m_width(o.m_width), m_height(o.m_height)
  {
    std::copy(o.m_data, o.m_data + (size_t)m_width * m_height * nChannels, m_data);
  }
  Image &operator=(const Image &o)
  {
    if (this == &o) return *this;
    if (!m_data || m_width != o.m_width || m_height != o.m_height) {
      delete[] m_data;
      m_data = new type[(size_t)o.m_width * o.m_height * nChannels];
      m_width = o.m_width;
      m_height = o.m_height;
    }
    std::copy(o.m_data, o.m_data + (size_t)m_width * m_height * nChannels, m_data);
    return *this;
  }
  ~Image() { delete[] m_data; }
  type *getData() const { return m_data; }
  unsigned int getWidth() const { return m_width; }
  unsigned int getHeight() const { return m_height; }
};
