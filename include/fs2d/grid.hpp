// fs2d/grid.hpp -- host grid with the layout of the reference's grid<T>
// (src/grid.hpp:8-66): rows x cols values, row-major, element (i, j) at
// i * cols + j, the one-cell boundary ring included.
//
// Same interface as the reference.  One deliberate difference: cols()
// returns the column count (the reference returns rows, src/grid.hpp:20-22;
// the solvers only accept square grids, where the two agree).
#pragma once

#include <cstddef>
#include <iterator>
#include <vector>

template <typename T>
class grid {
public:
    grid(std::size_t rows, std::size_t cols, T initial_value)
        : m_rows(rows), m_cols(cols), m_data(rows * cols, initial_value) {}

    std::size_t rows() const { return m_rows; }
    std::size_t cols() const { return m_cols; }

    const T* data() const { return m_data.data(); }
    T* data() { return m_data.data(); }

    T operator()(const std::size_t i, const std::size_t j) const { return m_data[i * m_cols + j]; }
    T& operator()(const std::size_t i, const std::size_t j) { return m_data[i * m_cols + j]; }

    auto begin() { return std::begin(m_data); }
    auto end() { return std::end(m_data); }
    auto cbegin() const { return std::cbegin(m_data); }
    auto cend() const { return std::cend(m_data); }

private:
    std::size_t m_rows;
    std::size_t m_cols;
    std::vector<T> m_data;
};
