// Complex arithmetic usable on host and device (HIP), layout (re, im) like
// C99 double complex / std::complex<double>.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

#define RQC_HD __host__ __device__ __forceinline__

template <typename T>
struct cx
{
	T x, y;  // real, imaginary
};

using zc = cx<double>;
using fc = cx<float>;

template <typename T> RQC_HD cx<T> mk(T re, T im) { return cx<T>{re, im}; }
template <typename T> RQC_HD cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> RQC_HD cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> RQC_HD cx<T> operator-(cx<T> a) { return {-a.x, -a.y}; }
template <typename T> RQC_HD cx<T> operator*(cx<T> a, cx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T> RQC_HD cx<T> operator*(T s, cx<T> a) { return {s * a.x, s * a.y}; }
template <typename T> RQC_HD cx<T>& operator+=(cx<T>& a, cx<T> b) { a.x += b.x; a.y += b.y; return a; }
template <typename T> RQC_HD cx<T>& operator-=(cx<T>& a, cx<T> b) { a.x -= b.x; a.y -= b.y; return a; }
template <typename T> RQC_HD cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
RQC_HD double cabs(zc a) { return hypot(a.x, a.y); }

// a / b with the textbook formula scaled by the larger component of b
RQC_HD zc cdiv(zc a, zc b)
{
	if (fabs(b.x) >= fabs(b.y)) {
		const double r = b.y / b.x, d = b.x + b.y * r;
		return {(a.x + a.y * r) / d, (a.y - a.x * r) / d};
	}
	const double r = b.x / b.y, d = b.x * r + b.y;
	return {(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}

// accumulate s += a * b with fused multiply-adds (4 FMA per complex MAC)
template <typename T>
RQC_HD void cmac(cx<T>& s, cx<T> a, cx<T> b)
{
	s.x = fma(a.x, b.x, s.x);
	s.x = fma(-a.y, b.y, s.x);
	s.y = fma(a.x, b.y, s.y);
	s.y = fma(a.y, b.x, s.y);
}

template <typename T>
RQC_HD cx<double> to_z(cx<T> a) { return {(double)a.x, (double)a.y}; }
