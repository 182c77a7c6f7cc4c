// H-LLt factorisation and triangular solves (placeholder until the task
// executor lands).
#include <stdexcept>

#include "engine.hpp"

namespace hmx {

struct FactorState {};

void Matrix::factorize(int, bool) { throw std::runtime_error("factorize: not implemented yet"); }
template <typename T>
void Matrix::solve_device(T*, int, int, const Store<T>&) {
  throw std::runtime_error("solve: not implemented yet");
}
template void Matrix::solve_device<double>(double*, int, int, const Store<double>&);
template void Matrix::solve_device<float>(float*, int, int, const Store<float>&);
void Matrix::solve_host(void*, int, bool) { throw std::runtime_error("solve: not implemented yet"); }
double Matrix::logdet() { throw std::runtime_error("logdet: not implemented yet"); }
int Matrix::solve_refined(double*, int, int, double, int, double*) {
  throw std::runtime_error("solve_refined: not implemented yet");
}

}  // namespace hmx
