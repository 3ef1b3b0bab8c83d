// Owns every argument of a sequence and hands out references to the subset
// an algorithm declares (reference interface: ArgumentManager.cuh:51-74).
#pragma once

#include <cstddef>
#include <tuple>

template<typename Tuple>
struct ArgumentManager {
  Tuple arguments_tuple;
  char* base_pointer = nullptr;

  void set_base_pointer(char* p) { base_pointer = p; }

  template<typename T>
  typename T::type* offset() const
  {
    return reinterpret_cast<typename T::type*>(std::get<T>(arguments_tuple).offset);
  }

  template<typename T>
  size_t size() const
  {
    return std::get<T>(arguments_tuple).size;
  }

  template<typename T>
  void set_offset(size_t byte_offset)
  {
    std::get<T>(arguments_tuple).offset = base_pointer + byte_offset;
  }

  template<typename T>
  void set_size(size_t count)
  {
    std::get<T>(arguments_tuple).size = count * sizeof(typename T::type);
  }
};

template<typename Arguments>
struct ArgumentRefManager;

template<typename... Arguments>
struct ArgumentRefManager<std::tuple<Arguments...>> {
  std::tuple<Arguments&...> m_arguments;

  explicit ArgumentRefManager(std::tuple<Arguments&...> refs) : m_arguments(refs) {}

  template<typename T>
  typename T::type* offset() const
  {
    return reinterpret_cast<typename T::type*>(std::get<T&>(m_arguments).offset);
  }

  template<typename T>
  size_t size() const
  {
    return std::get<T&>(m_arguments).size;
  }

  template<typename T>
  void set_size(size_t count)
  {
    std::get<T&>(m_arguments).size = count * sizeof(typename T::type);
  }

  // Binds T to caller-owned device memory instead of the arena.
  template<typename T>
  void set_external(void* pointer, size_t bytes)
  {
    auto& a = std::get<T&>(m_arguments);
    a.offset = static_cast<char*>(pointer);
    a.size = bytes;
    a.external = true;
  }

  template<typename T>
  void clear_external()
  {
    std::get<T&>(m_arguments).external = false;
  }
};
