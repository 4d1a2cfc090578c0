// tg_ewald_table.hpp -- the periodic correction (tigergrav/green.hpp:189-286) as a table.
//
// The correction phi_c(x) = sum_{n != 0} -erfc(2|x - n|)/|x - n| + erf(2r)/r
//                           - sum_{h != 0} exp(-pi^2 h^2 / 4) / (pi h^2) cos(2 pi h.x) + pi/4
// (the Ewald sum of -1/r minus its nearest image; real space |n|^2 <= 17, Fourier |h|^2 <= 9, as
// tigergrav/green.hpp:3-59 truncates them) is smooth over the wrapped cell |x_i| <= 1/2, even in
// each coordinate and symmetric under permutations of the coordinates.  Its derivatives to
// seventh order are tabulated in FP64, rounded to float, on the points of the grid
// {0, 1/64, ..., 1/2}^3 with i >= j >= l (6545 points).  A row stores the 64 slots that hold the
// third axis at most once (1.8 MB in all: it stays in one XCD's L2); the correction's Laplacian
// is constant (-4 pi), so the other slots follow from the trace relation.  A kernel sorts |x| into
// that wedge, takes the derivatives it needs from a third-order Taylor expansion about the
// nearest point, permutes the components back and flips the sign of every component with an odd
// count of a negative coordinate's index.
//
// Symmetric tensors are stored by slot: orders 0..7 in turn, each order's sorted index tuples in
// lexicographic order (orders 0..4 are the expansion layout of tigergrav/expansion.hpp:77-137).
#pragma once
#include <stdint.h>

#include <vector>

#define TG_EWT_N 33    // grid points per axis over [0, 1/2]
#define TG_EWT_INV 64  // inverse grid spacing
#define TG_EWT_ORDER 7  // highest derivative order tabulated
#define TG_EWT_S 120    // slots of orders 0..7
#define TG_EWT_P 6545   // grid points with i >= j >= l: TG_EWT_N (TG_EWT_N + 1) (TG_EWT_N + 2) / 6
#define TG_EWT_ROWS (TG_EWT_P + 1)  // and one row for coincident points: the reference's constants
#define TG_EWT_C 64      // slots stored per row: those holding the third axis at most once (2 n + 1 of order n)
#define TG_EWT_STRIDE 68  // floats per row: 272 bytes

namespace tg {
namespace ewt {

struct Cnt {
	int n0, n1, n2;
};
struct SlotTable {
	Cnt c[TG_EWT_S];
};
constexpr SlotTable make_slots() {
	SlotTable t{};
	int k = 0;
	for (int n = 0; n <= TG_EWT_ORDER; n++)
		for (int n0 = n; n0 >= 0; n0--)
			for (int n1 = n - n0; n1 >= 0; n1--) {
				t.c[k].n0 = n0;
				t.c[k].n1 = n1;
				t.c[k].n2 = n - n0 - n1;
				k++;
			}
	return t;
}
constexpr SlotTable ST = make_slots();
constexpr int slot_of(int n0, int n1, int n2) {
	for (int k = 0; k < TG_EWT_S; k++)
		if (ST.c[k].n0 == n0 && ST.c[k].n1 == n1 && ST.c[k].n2 == n2) return k;
	return -1;
}
// slot k with indices a and b exchanged (a transposition of the coordinates)
constexpr int swap_ab(int k, int a, int b) {
	int n[3] = {ST.c[k].n0, ST.c[k].n1, ST.c[k].n2};
	const int t = n[a];
	n[a] = n[b];
	n[b] = t;
	return slot_of(n[0], n[1], n[2]);
}
// table row of the sorted grid point i >= j >= l
constexpr int wedge_index(int i, int j, int l) { return i * (i + 1) * (i + 2) / 6 + j * (j + 1) / 2 + l; }
static_assert(wedge_index(TG_EWT_N - 1, TG_EWT_N - 1, TG_EWT_N - 1) == TG_EWT_P - 1, "wedge size");
// a row's position of slot k (stored slots in slot order), or -1 for a slot the trace relation
// gives: the correction's Laplacian is constant (-4 pi), so D_{k} = C - D_{k - 2z + 2x} - D_{k - 2z + 2y}
// for every slot holding the third axis twice or more (C = -4 pi for D_zz, else 0)
constexpr int stored_index(int k) {
	if (ST.c[k].n2 > 1) return -1;
	int c = 0;
	for (int i = 0; i < k; i++) c += ST.c[i].n2 <= 1 ? 1 : 0;
	return c;
}
// stored slots of orders 0..m
constexpr int stored_count(int m) { return (m + 1) * (m + 1); }
static_assert(stored_count(TG_EWT_ORDER) == TG_EWT_C, "stored slots");
static_assert(TG_EWT_STRIDE >= TG_EWT_C && TG_EWT_STRIDE % 4 == 0, "row stride");
// slot k with one more index a, with indices a and b
constexpr int up1(int k, int a) { return slot_of(ST.c[k].n0 + (a == 0), ST.c[k].n1 + (a == 1), ST.c[k].n2 + (a == 2)); }
constexpr int up2(int k, int a, int b) { return up1(up1(k, a), b); }
// slot of the multiset union of slots k and j
constexpr int merge(int k, int j) { return slot_of(ST.c[k].n0 + ST.c[j].n0, ST.c[k].n1 + ST.c[j].n1, ST.c[k].n2 + ST.c[j].n2); }
// whether component k changes sign with coordinate a: an odd count of index a
constexpr bool odd_in(int k, int a) { return ((a == 0 ? ST.c[k].n0 : a == 1 ? ST.c[k].n1 : ST.c[k].n2) & 1) != 0; }
// 1 / (j0! j1! j2!) of slot j: the Taylor coefficient of the monomial d^j
constexpr float taylor_coef(int j) {
	float f = 1.0f;
	for (int i = 2; i <= ST.c[j].n0; i++) f *= float(i);
	for (int i = 2; i <= ST.c[j].n1; i++) f *= float(i);
	for (int i = 2; i <= ST.c[j].n2; i++) f *= float(i);
	return 1.0f / f;
}
static_assert(slot_of(2, 0, 0) == 4 && slot_of(0, 1, 1) == 8 && slot_of(1, 1, 1) == 14 && slot_of(0, 0, 3) == 19 &&
				slot_of(2, 1, 1) == 24 && slot_of(0, 0, 4) == 34 && slot_of(6, 0, 0) == 56 && slot_of(0, 0, 7) == 119,
		"slot layout");

}  // namespace ewt

// FP64 value of the correction and its TG_EWT_S derivatives (orders 0..TG_EWT_ORDER) at x (host);
// at x = 0 the analytic limit (the kernels apply the reference's r -> 0 convention themselves)
void ewald_green64(const double* x, double* D);
// the table: TG_EWT_ROWS rows of TG_EWT_STRIDE floats (the TG_EWT_C stored slots in slot order, then
// zeros): the sorted grid points (row
// wedge_index(i, j, l)), then row TG_EWT_P = (2.8372975, 0, ..., 0), the reference's values at
// r -> 0, which a kernel reads with a zero Taylor offset for a coincident pair
void ewald_table_build(std::vector<float>& T);

}  // namespace tg
