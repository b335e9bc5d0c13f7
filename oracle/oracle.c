/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * CPU restatement of the reference's gate-sweep path.  Every function names
 * the reference file:line range whose algorithm (and, for parity, whose
 * summation order) it follows.  Built with plain gcc -O3 (no -ffast-math, no
 * FMA), like the reference itself.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static void* xalloc(size_t n)
{
	void* p = malloc(n > 0 ? n : 1);
	if (p == NULL) {
		fprintf(stderr, "oracle: allocation of %zu bytes failed\n", n);
		abort();
	}
	return p;
}

/* exchange the roles of the two wires of a gate (rows and columns of |q_i q_j>):
   reference src/gate.c:17-27 */
static void swap_gate_wires(const ocplx* g, ocplx* gp)
{
	static const int p[4] = { 0, 2, 1, 3 };
	for (int r = 0; r < 4; r++) {
		for (int c = 0; c < 4; c++) {
			gp[4*r + c] = g[4*p[r] + p[c]];
		}
	}
}

/* base index of quad number q for qubits i < j: zero bits inserted at
   positions L-1-i (high) and L-1-j (low) */
static inline long quad_base(long q, int L, int i, int j)
{
	const int lo = L - 1 - j, hi = L - 1 - i;
	long x = ((q >> lo) << (lo + 1)) | (q & ((1L << lo) - 1));
	x = ((x >> hi) << (hi + 1)) | (x & ((1L << hi) - 1));
	return x;
}

/* ---------------------------------------------------------------------------
 * Gate application, reference src/gate.c:10-76.  For j < i the gate is
 * re-indexed and applied with the wires exchanged (:17-30).  Quads are visited
 * in increasing base index, matching the (a, b, c) loop nest (:31-75).
 */
void orc_apply_gate(const ocplx* gate, int i, int j, int L, const ocplx* psi, ocplx* psi_out)
{
	if (j < i) {
		ocplx gp[16];
		swap_gate_wires(gate, gp);
		orc_apply_gate(gp, j, i, L, psi, psi_out);
		return;
	}
	const long nq = 1L << (L - 2);
	const long si = 1L << (L - 1 - i), sj = 1L << (L - 1 - j);
	for (long q = 0; q < nq; q++) {
		const long k = quad_base(q, L, i, j);
		const ocplx x = psi[k], y = psi[k + sj], z = psi[k + si], w = psi[k + si + sj];
		psi_out[k]           = gate[ 0]*x + gate[ 1]*y + gate[ 2]*z + gate[ 3]*w;
		psi_out[k + sj]      = gate[ 4]*x + gate[ 5]*y + gate[ 6]*z + gate[ 7]*w;
		psi_out[k + si]      = gate[ 8]*x + gate[ 9]*y + gate[10]*z + gate[11]*w;
		psi_out[k + si + sj] = gate[12]*x + gate[13]*y + gate[14]*z + gate[15]*w;
	}
}

/* ---------------------------------------------------------------------------
 * Backward pass of one gate, reference src/gate.c:83-188:
 *   dpsi = gate^T dpsi_out;  dgate = sum over quads of dpsi_out (x) psi
 * (holomorphic convention, no conjugation).  dgate is zeroed and then
 * accumulated quad by quad; for j < i it is un-permuted at the end (:110-113).
 */
void orc_apply_gate_backward(const ocplx* gate, int i, int j, int L, const ocplx* psi,
	const ocplx* dpsi_out, ocplx* dpsi, ocplx* dgate)
{
	if (j < i) {
		ocplx gp[16], dgp[16];
		swap_gate_wires(gate, gp);
		orc_apply_gate_backward(gp, j, i, L, psi, dpsi_out, dpsi, dgp);
		swap_gate_wires(dgp, dgate);
		return;
	}
	for (int e = 0; e < 16; e++) { dgate[e] = 0; }
	const long nq = 1L << (L - 2);
	const long si = 1L << (L - 1 - i), sj = 1L << (L - 1 - j);
	for (long q = 0; q < nq; q++) {
		const long k = quad_base(q, L, i, j);
		const long idx[4] = { k, k + sj, k + si, k + si + sj };
		ocplx v[4], d[4];
		for (int t = 0; t < 4; t++) { v[t] = psi[idx[t]]; d[t] = dpsi_out[idx[t]]; }
		for (int r = 0; r < 4; r++) {
			for (int c = 0; c < 4; c++) {
				dgate[4*r + c] += d[r] * v[c];
			}
		}
		for (int c = 0; c < 4; c++) {
			dpsi[idx[c]] = gate[c]*d[0] + gate[4 + c]*d[1] + gate[8 + c]*d[2] + gate[12 + c]*d[3];
		}
	}
}

/* Single-qubit gate on wire i (qubit i = bit L-1-i).  The reference has two-qubit gates
   only; this follows its two-qubit conventions (src/gate.c:10-76 and :83-188: row-major
   gate, dpsi = gate^T dpsi_out, dgate = sum dpsi_out (x) psi without conjugation), pair by
   pair in ascending order.  Parity unpinned by the reference: pinned by dense Kronecker
   products and the bilinear identity in tests/test_oracle_pinning.py. */
void orc_apply_gate_1q(const ocplx* gate, int i, int L, const ocplx* psi, ocplx* psi_out)
{
	const long np = 1L << (L - 1);
	const int b = L - 1 - i;
	const long s = 1L << b;
	for (long q = 0; q < np; q++) {
		const long k = ((q >> b) << (b + 1)) | (q & (s - 1));
		const ocplx x = psi[k], y = psi[k + s];
		psi_out[k]     = gate[0]*x + gate[1]*y;
		psi_out[k + s] = gate[2]*x + gate[3]*y;
	}
}

void orc_apply_gate_1q_backward(const ocplx* gate, int i, int L, const ocplx* psi, const ocplx* dpsi_out,
	ocplx* dpsi, ocplx* dgate)
{
	const long np = 1L << (L - 1);
	const int b = L - 1 - i;
	const long s = 1L << b;
	for (int e = 0; e < 4; e++) { dgate[e] = 0; }
	for (long q = 0; q < np; q++) {
		const long k = ((q >> b) << (b + 1)) | (q & (s - 1));
		const ocplx v[2] = { psi[k], psi[k + s] };
		const ocplx d[2] = { dpsi_out[k], dpsi_out[k + s] };
		for (int r = 0; r < 2; r++) {
			for (int c = 0; c < 2; c++) { dgate[2*r + c] += d[r] * v[c]; }
		}
		dpsi[k]     = gate[0]*d[0] + gate[2]*d[1];
		dpsi[k + s] = gate[1]*d[0] + gate[3]*d[1];
	}
}

/* Gate on an interleaved statevector array, reference src/gate.c:195-262:
   the trailing stride is multiplied by nstates. */
void orc_apply_gate_to_array(const ocplx* gate, int i, int j, int L, int nstates, const ocplx* psi, ocplx* psi_out)
{
	if (j < i) {
		ocplx gp[16];
		swap_gate_wires(gate, gp);
		orc_apply_gate_to_array(gp, j, i, L, nstates, psi, psi_out);
		return;
	}
	const long nq = 1L << (L - 2);
	const long si = (1L << (L - 1 - i)) * nstates, sj = (1L << (L - 1 - j)) * nstates;
	for (long q = 0; q < nq; q++) {
		const long kb = quad_base(q, L, i, j) * nstates;
		for (int s = 0; s < nstates; s++) {
			const long k = kb + s;
			const ocplx x = psi[k], y = psi[k + sj], z = psi[k + si], w = psi[k + si + sj];
			psi_out[k]           = gate[ 0]*x + gate[ 1]*y + gate[ 2]*z + gate[ 3]*w;
			psi_out[k + sj]      = gate[ 4]*x + gate[ 5]*y + gate[ 6]*z + gate[ 7]*w;
			psi_out[k + si]      = gate[ 8]*x + gate[ 9]*y + gate[10]*z + gate[11]*w;
			psi_out[k + si + sj] = gate[12]*x + gate[13]*y + gate[14]*z + gate[15]*w;
		}
	}
}

/* ---------------------------------------------------------------------------
 * Circuits of two-qubit gates on explicit wires.
 */

/* reference src/quantum_circuit.c:13-35 (sequential application) */
int orc_apply_quantum_circuit(const ocplx* gates, int ngates, const int* wires, int L, const ocplx* psi, ocplx* psi_out)
{
	const long n = 1L << L;
	ocplx* a = xalloc(n * sizeof(ocplx));
	ocplx* b = xalloc(n * sizeof(ocplx));
	memcpy(a, psi, n * sizeof(ocplx));
	for (int g = 0; g < ngates; g++) {
		orc_apply_gate(&gates[16*g], wires[2*g], wires[2*g + 1], L, a, b);
		ocplx* t = a; a = b; b = t;
	}
	memcpy(psi_out, a, n * sizeof(ocplx));
	free(a); free(b);
	return 0;
}

/* reference src/quantum_circuit.c:90-108: cache[g] = state before gate g */
int orc_quantum_circuit_forward(const ocplx* gates, int ngates, const int* wires, int L, const ocplx* psi,
	ocplx* cache, ocplx* psi_out)
{
	const long n = 1L << L;
	memcpy(cache, psi, n * sizeof(ocplx));
	for (int g = 0; g < ngates; g++) {
		orc_apply_gate(&gates[16*g], wires[2*g], wires[2*g + 1], L, &cache[g*n],
			g + 1 < ngates ? &cache[(g + 1)*n] : psi_out);
	}
	return 0;
}

/* reference src/quantum_circuit.c:115-140: reverse sweep, dgates[g] per gate */
int orc_quantum_circuit_backward(const ocplx* gates, int ngates, const int* wires, int L, const ocplx* cache,
	const ocplx* dpsi_out, ocplx* dpsi, ocplx* dgates)
{
	const long n = 1L << L;
	ocplx* cur = xalloc(n * sizeof(ocplx));
	ocplx* nxt = xalloc(n * sizeof(ocplx));
	memcpy(cur, dpsi_out, n * sizeof(ocplx));
	for (int g = ngates - 1; g >= 0; g--) {
		orc_apply_gate_backward(&gates[16*g], wires[2*g], wires[2*g + 1], L, &cache[g*n], cur, nxt, &dgates[16*g]);
		ocplx* t = cur; cur = nxt; nxt = t;
	}
	memcpy(dpsi, cur, n * sizeof(ocplx));
	free(cur); free(nxt);
	return 0;
}

/* reference src/quantum_circuit.c:147-236.  Steps:
   (1) forward cache psi_g (state before gate g);
   (2) reverse sweep from phi storing phi_g (adjoint after gate g) and dgates;
   (3) tangent-forward  dpsi_{g+1} = G_g dpsi_g + Z_g psi_g, hess_g  = phi_g (x) dpsi_g;
   (4) tangent-backward dphi_{g-1} = G_g^T dphi_g + Z_g^T phi_g, hess_g += dphi_g (x) psi_g.
   Gate-level calls and their operand order follow the reference exactly. */
int orc_quantum_circuit_gates_hessian_vector_product(const ocplx* gates, const ocplx* gatedirs, int ngates,
	const int* wires, int L, const ocplx* psi, const ocplx* phi, ocplx* psi_out, ocplx* dgates, ocplx* hess)
{
	const long n = 1L << L;
	ocplx* cache = xalloc((size_t)ngates * n * sizeof(ocplx));
	ocplx* cache_rev = xalloc((size_t)ngates * n * sizeof(ocplx));
	ocplx* tmp1 = xalloc(n * sizeof(ocplx));
	ocplx* tmp2 = xalloc(n * sizeof(ocplx));
	ocplx* dv = xalloc(n * sizeof(ocplx));

	orc_quantum_circuit_forward(gates, ngates, wires, L, psi, cache, psi_out);

	memcpy(&cache_rev[(ngates - 1)*n], phi, n * sizeof(ocplx));
	for (int g = ngates - 1; g >= 0; g--) {
		ocplx* out = g > 0 ? &cache_rev[(g - 1)*n] : tmp1;  /* phi_out of the last step is discarded */
		orc_apply_gate_backward(&gates[16*g], wires[2*g], wires[2*g + 1], L, &cache[g*n], &cache_rev[g*n], out, &dgates[16*g]);
	}

	/* tangent forward, :183-204 */
	for (int e = 0; e < 16; e++) { hess[e] = 0; }
	orc_apply_gate(&gatedirs[0], wires[0], wires[1], L, &cache[0], dv);
	for (int g = 1; g < ngates; g++) {
		ocplx gt[16], dg[16];
		orc_transpose(&gates[16*g], gt);
		orc_apply_gate_backward(gt, wires[2*g], wires[2*g + 1], L, &cache_rev[g*n], dv, tmp1, dg);
		orc_transpose(dg, &hess[16*g]);
		orc_apply_gate(&gatedirs[16*g], wires[2*g], wires[2*g + 1], L, &cache[g*n], tmp2);
		for (long a = 0; a < n; a++) { dv[a] = tmp1[a] + tmp2[a]; }
	}

	/* tangent backward, :207-226 */
	{
		const int g = ngates - 1;
		ocplx zt[16];
		orc_transpose(&gatedirs[16*g], zt);
		orc_apply_gate(zt, wires[2*g], wires[2*g + 1], L, &cache_rev[g*n], dv);
	}
	for (int g = ngates - 2; g >= 0; g--) {
		ocplx dg[16], zt[16];
		orc_apply_gate_backward(&gates[16*g], wires[2*g], wires[2*g + 1], L, &cache[g*n], dv, tmp1, dg);
		for (int e = 0; e < 16; e++) { hess[16*g + e] += dg[e]; }
		orc_transpose(&gatedirs[16*g], zt);
		orc_apply_gate(zt, wires[2*g], wires[2*g + 1], L, &cache_rev[g*n], tmp2);
		for (long a = 0; a < n; a++) { dv[a] = tmp1[a] + tmp2[a]; }
	}

	free(dv); free(tmp2); free(tmp1); free(cache_rev); free(cache);
	return 0;
}

/* ---------------------------------------------------------------------------
 * Brickwall circuits: layer i applies vlist[i] to the pairs
 * (perms[i][2k], perms[i][2k+1]), k = 0 .. L/2-1.
 */

/* expansion into a sequential gate list, reference src/target.c:443-458 */
static void brickwall_to_sequential(int L, int nlayers, const ocplx* vlist, const int* perms, ocplx* gates, int* wires)
{
	const int per = L / 2;
	for (int i = 0; i < nlayers; i++) {
		for (int k = 0; k < per; k++) {
			memcpy(&gates[16*(i*per + k)], &vlist[16*i], 16 * sizeof(ocplx));
			wires[2*(i*per + k)]     = perms[i*L + 2*k];
			wires[2*(i*per + k) + 1] = perms[i*L + 2*k + 1];
		}
	}
}

/* reference src/brickwall_circuit.c:14-44 */
int orc_apply_brickwall_unitary(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* psi, ocplx* psi_out)
{
	const int ng = nlayers * (L / 2);
	ocplx* gates = xalloc((size_t)ng * 16 * sizeof(ocplx));
	int* wires = xalloc((size_t)ng * 2 * sizeof(int));
	brickwall_to_sequential(L, nlayers, vlist, perms, gates, wires);
	int ret = orc_apply_quantum_circuit(gates, ng, wires, L, psi, psi_out);
	free(wires); free(gates);
	return ret;
}

/* reference src/brickwall_circuit.c:51-78 */
int orc_brickwall_unitary_forward(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* psi,
	ocplx* cache, ocplx* psi_out)
{
	const int ng = nlayers * (L / 2);
	ocplx* gates = xalloc((size_t)ng * 16 * sizeof(ocplx));
	int* wires = xalloc((size_t)ng * 2 * sizeof(int));
	brickwall_to_sequential(L, nlayers, vlist, perms, gates, wires);
	int ret = orc_quantum_circuit_forward(gates, ng, wires, L, psi, cache, psi_out);
	free(wires); free(gates);
	return ret;
}

/* reference src/brickwall_circuit.c:85-124: per layer, gradients of its gates
   are summed in reverse pair order (k = L/2-1 down to 0) */
int orc_brickwall_unitary_backward(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* cache,
	const ocplx* dpsi_out, ocplx* dpsi, ocplx* dvlist)
{
	const long n = 1L << L;
	const int per = L / 2;
	ocplx* cur = xalloc(n * sizeof(ocplx));
	ocplx* nxt = xalloc(n * sizeof(ocplx));
	memcpy(cur, dpsi_out, n * sizeof(ocplx));
	int g = nlayers * per - 1;
	for (int i = nlayers - 1; i >= 0; i--) {
		for (int e = 0; e < 16; e++) { dvlist[16*i + e] = 0; }
		for (int k = per - 1; k >= 0; k--) {
			ocplx dV[16];
			orc_apply_gate_backward(&vlist[16*i], perms[i*L + 2*k], perms[i*L + 2*k + 1], L, &cache[g*n], cur, nxt, dV);
			for (int e = 0; e < 16; e++) { dvlist[16*i + e] += dV[e]; }
			ocplx* t = cur; cur = nxt; nxt = t;
			g--;
		}
	}
	memcpy(dpsi, cur, n * sizeof(ocplx));
	free(cur); free(nxt);
	return 0;
}

/* ---------------------------------------------------------------------------
 * Targets f = -Tr(U^dagger C) via a trace over basis columns.
 */

int orc_ufunc_dense(int nqubits, const ocplx* psi, void* udata, ocplx* psi_out)
{
	const long n = 1L << nqubits;
	const ocplx* U = udata;
	for (long a = 0; a < n; a++) {
		ocplx s = 0;
		for (long b = 0; b < n; b++) { s += U[a*n + b] * psi[b]; }
		psi_out[a] = s;
	}
	return 0;
}

/* the index-scrambling linear map of the reference tests, test/test_target.c:16-34 */
int orc_ufunc_scramble(int nqubits, const ocplx* psi, void* udata, ocplx* psi_out)
{
	(void)udata;
	const long n = 1L << nqubits;
	for (long i = 0; i < n; i++) {
		psi_out[i] = (-1.1 + 0.8*I) * psi[((i + 3) * 113) % n] + (0.4 - 0.7*I) * psi[((i + 9) * 173) % n]
		           + (0.5 + 0.1*I) * psi[i] + (-0.3 + 0.2*I) * psi[((i + 4) * 199) % n];
	}
	return 0;
}

int orc_ufunc_brickwall(int nqubits, const ocplx* psi, void* udata, ocplx* psi_out)
{
	const struct orc_brickwall_op* op = udata;
	return orc_apply_brickwall_unitary(op->vlist, op->nlayers, op->perms, nqubits, psi, psi_out);
}

/* evaluates -conj(U e_b) into upsi, reference src/target.c:37-54 */
static int adjoint_seed(orc_linear_func ufunc, void* udata, int L, long b, ocplx* psi, ocplx* upsi)
{
	const long n = 1L << L;
	memset(psi, 0, n * sizeof(ocplx));
	psi[b] = 1;
	int ret = ufunc(L, psi, udata, upsi);
	if (ret < 0) { return -2; }
	for (long a = 0; a < n; a++) { upsi[a] = -conj(upsi[a]); }
	return 0;
}

/* reference src/target.c:15-75 */
int orc_circuit_unitary_target(orc_linear_func ufunc, void* udata, const ocplx* gates, int ngates, const int* wires,
	int L, ocplx* fval)
{
	const long n = 1L << L;
	ocplx* psi = xalloc(n * sizeof(ocplx));
	ocplx* upsi = xalloc(n * sizeof(ocplx));
	ocplx* cpsi = xalloc(n * sizeof(ocplx));
	ocplx f = 0;
	for (long b = 0; b < n; b++) {
		if (adjoint_seed(ufunc, udata, L, b, psi, upsi) < 0) { return -2; }
		orc_apply_quantum_circuit(gates, ngates, wires, L, psi, cpsi);
		for (long a = 0; a < n; a++) { f += upsi[a] * cpsi[a]; }
	}
	free(cpsi); free(upsi); free(psi);
	*fval = f;
	return 0;
}

/* reference src/target.c:158-247 */
int orc_circuit_unitary_target_and_gradient(orc_linear_func ufunc, void* udata, const ocplx* gates, int ngates,
	const int* wires, int L, ocplx* fval, ocplx* dgates)
{
	const long n = 1L << L;
	ocplx* psi = xalloc(n * sizeof(ocplx));
	ocplx* upsi = xalloc(n * sizeof(ocplx));
	ocplx* cpsi = xalloc(n * sizeof(ocplx));
	ocplx* cache = xalloc((size_t)ngates * n * sizeof(ocplx));
	ocplx* dunit = xalloc((size_t)ngates * 16 * sizeof(ocplx));
	for (int e = 0; e < 16*ngates; e++) { dgates[e] = 0; }
	ocplx f = 0;
	for (long b = 0; b < n; b++) {
		if (adjoint_seed(ufunc, udata, L, b, psi, upsi) < 0) { return -2; }
		orc_quantum_circuit_forward(gates, ngates, wires, L, psi, cache, cpsi);
		for (long a = 0; a < n; a++) { f += upsi[a] * cpsi[a]; }
		orc_quantum_circuit_backward(gates, ngates, wires, L, cache, upsi, psi, dunit);
		for (int e = 0; e < 16*ngates; e++) { dgates[e] += dunit[e]; }
	}
	free(dunit); free(cache); free(cpsi); free(upsi); free(psi);
	*fval = f;
	return 0;
}

/* reference src/target.c:256-280 */
int orc_circuit_unitary_target_and_projected_gradient(orc_linear_func ufunc, void* udata, const ocplx* gates,
	int ngates, const int* wires, int L, ocplx* fval, double* grad_vec)
{
	ocplx* dgates = xalloc((size_t)ngates * 16 * sizeof(ocplx));
	int ret = orc_circuit_unitary_target_and_gradient(ufunc, udata, gates, ngates, wires, L, fval, dgates);
	if (ret < 0) { free(dgates); return ret; }
	for (int g = 0; g < ngates; g++) {
		for (int e = 0; e < 16; e++) { dgates[16*g + e] = conj(dgates[16*g + e]); }
		orc_tangent_to_real(&gates[16*g], &dgates[16*g], &grad_vec[16*g]);
	}
	free(dgates);
	return 0;
}

/* column range [b0, b1) of the HVP trace, reference src/target.c:335-369 */
static int circuit_hvp_columns(orc_linear_func ufunc, void* udata, const ocplx* gates, const ocplx* gatedirs,
	int ngates, const int* wires, int L, long b0, long b1, ocplx* fval, ocplx* dgates, ocplx* hess)
{
	const long n = 1L << L;
	ocplx* psi = xalloc(n * sizeof(ocplx));
	ocplx* upsi = xalloc(n * sizeof(ocplx));
	ocplx* cpsi = xalloc(n * sizeof(ocplx));
	ocplx* dunit = xalloc((size_t)ngates * 16 * sizeof(ocplx));
	ocplx* hunit = xalloc((size_t)ngates * 16 * sizeof(ocplx));
	for (int e = 0; e < 16*ngates; e++) { dgates[e] = 0; hess[e] = 0; }
	ocplx f = 0;
	for (long b = b0; b < b1; b++) {
		if (adjoint_seed(ufunc, udata, L, b, psi, upsi) < 0) { return -2; }
		orc_quantum_circuit_gates_hessian_vector_product(gates, gatedirs, ngates, wires, L, psi, upsi, cpsi, dunit, hunit);
		for (long a = 0; a < n; a++) { f += upsi[a] * cpsi[a]; }
		for (int e = 0; e < 16*ngates; e++) {
			dgates[e] += dunit[e];
			hess[e] += hunit[e];
		}
	}
	free(hunit); free(dunit); free(cpsi); free(upsi); free(psi);
	*fval = f;
	return 0;
}

/* reference src/target.c:289-380 */
int orc_circuit_unitary_target_hessian_vector_product(orc_linear_func ufunc, void* udata, const ocplx* gates,
	const ocplx* gatedirs, int ngates, const int* wires, int L, ocplx* fval, ocplx* dgates, ocplx* hess)
{
	return circuit_hvp_columns(ufunc, udata, gates, gatedirs, ngates, wires, L, 0, 1L << L, fval, dgates, hess);
}

/* Riemannian projection of gradient and HVP for one gate,
   reference src/target.c:404-430 and :650-671 */
static void project_grad_hvp(const ocplx* v, const ocplx* z, ocplx* dv, ocplx* hz, double* grad, double* hvp)
{
	for (int e = 0; e < 16; e++) { dv[e] = conj(dv[e]); }
	orc_tangent_to_real(v, dv, grad);
	for (int e = 0; e < 16; e++) { hz[e] = conj(hz[e]); }
	ocplx gradh[16], pderiv[16];
	orc_adjoint(dv, gradh);
	orc_symmetric_triple_matrix_product(z, gradh, v, pderiv);
	for (int e = 0; e < 16; e++) { hz[e] -= pderiv[e]; }
	orc_tangent_to_real(v, hz, hvp);
}

/* reference src/target.c:389-436 */
int orc_circuit_unitary_target_projected_hessian_vector_product(orc_linear_func ufunc, void* udata,
	const ocplx* gates, const ocplx* gatedirs, int ngates, const int* wires, int L, ocplx* fval,
	double* grad_vec, double* hvp_vec)
{
	ocplx* dgates = xalloc((size_t)ngates * 16 * sizeof(ocplx));
	ocplx* hess = xalloc((size_t)ngates * 16 * sizeof(ocplx));
	int ret = orc_circuit_unitary_target_hessian_vector_product(ufunc, udata, gates, gatedirs, ngates, wires, L, fval, dgates, hess);
	if (ret < 0) { free(hess); free(dgates); return ret; }
	for (int g = 0; g < ngates; g++) {
		project_grad_hvp(&gates[16*g], &gatedirs[16*g], &dgates[16*g], &hess[16*g], &grad_vec[16*g], &hvp_vec[16*g]);
	}
	free(hess); free(dgates);
	return 0;
}

/* reference src/target.c:467-481 */
int orc_brickwall_unitary_target(orc_linear_func ufunc, void* udata, const ocplx* vlist, int nlayers, int L,
	const int* perms, ocplx* fval)
{
	const int ng = nlayers * (L / 2);
	ocplx* gates = xalloc((size_t)ng * 16 * sizeof(ocplx));
	int* wires = xalloc((size_t)ng * 2 * sizeof(int));
	brickwall_to_sequential(L, nlayers, vlist, perms, gates, wires);
	int ret = orc_circuit_unitary_target(ufunc, udata, gates, ng, wires, L, fval);
	free(wires); free(gates);
	return ret;
}

/* per-layer sum of the per-gate matrices in pair order, reference src/target.c:526-535 */
static void sum_layers(const ocplx* per_gate, int nlayers, int per, ocplx* per_layer)
{
	for (int i = 0; i < nlayers; i++) {
		for (int e = 0; e < 16; e++) { per_layer[16*i + e] = 0; }
		for (int k = 0; k < per; k++) {
			for (int e = 0; e < 16; e++) { per_layer[16*i + e] += per_gate[16*(i*per + k) + e]; }
		}
	}
}

/* reference src/target.c:514-542 */
int orc_brickwall_unitary_target_and_gradient(orc_linear_func ufunc, void* udata, const ocplx* vlist, int nlayers,
	int L, const int* perms, ocplx* fval, ocplx* dvlist)
{
	const int ng = nlayers * (L / 2);
	ocplx* gates = xalloc((size_t)ng * 16 * sizeof(ocplx));
	ocplx* dgates = xalloc((size_t)ng * 16 * sizeof(ocplx));
	int* wires = xalloc((size_t)ng * 2 * sizeof(int));
	brickwall_to_sequential(L, nlayers, vlist, perms, gates, wires);
	int ret = orc_circuit_unitary_target_and_gradient(ufunc, udata, gates, ng, wires, L, fval, dgates);
	sum_layers(dgates, nlayers, L / 2, dvlist);
	free(wires); free(dgates); free(gates);
	return ret;
}

/* reference src/target.c:551-575 */
int orc_brickwall_unitary_target_and_projected_gradient(orc_linear_func ufunc, void* udata, const ocplx* vlist,
	int nlayers, int L, const int* perms, ocplx* fval, double* grad_vec)
{
	ocplx* dvlist = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	int ret = orc_brickwall_unitary_target_and_gradient(ufunc, udata, vlist, nlayers, L, perms, fval, dvlist);
	if (ret < 0) { free(dvlist); return ret; }
	for (int i = 0; i < nlayers; i++) {
		for (int e = 0; e < 16; e++) { dvlist[16*i + e] = conj(dvlist[16*i + e]); }
		orc_tangent_to_real(&vlist[16*i], &dvlist[16*i], &grad_vec[16*i]);
	}
	free(dvlist);
	return 0;
}

int orc_brickwall_hvp_columns(orc_linear_func ufunc, void* udata, const ocplx* vlist, const ocplx* vdirs,
	int nlayers, const int* perms, int L, long b0, long b1, ocplx* fval, ocplx* dvlist, ocplx* hess_vdirs)
{
	const int ng = nlayers * (L / 2);
	ocplx* gates = xalloc((size_t)ng * 16 * sizeof(ocplx));
	ocplx* gdirs = xalloc((size_t)ng * 16 * sizeof(ocplx));
	ocplx* dgates = xalloc((size_t)ng * 16 * sizeof(ocplx));
	ocplx* hess = xalloc((size_t)ng * 16 * sizeof(ocplx));
	int* wires = xalloc((size_t)ng * 2 * sizeof(int));
	brickwall_to_sequential(L, nlayers, vlist, perms, gates, wires);
	brickwall_to_sequential(L, nlayers, vdirs, perms, gdirs, wires);
	int ret = circuit_hvp_columns(ufunc, udata, gates, gdirs, ng, wires, L, b0, b1, fval, dgates, hess);
	sum_layers(dgates, nlayers, L / 2, dvlist);
	sum_layers(hess, nlayers, L / 2, hess_vdirs);
	free(wires); free(hess); free(dgates); free(gdirs); free(gates);
	return ret;
}

/* reference src/target.c:584-622 */
int orc_brickwall_unitary_target_hessian_vector_product(orc_linear_func ufunc, void* udata, const ocplx* vlist,
	const ocplx* vdirs, int nlayers, const int* perms, int L, ocplx* fval, ocplx* dvlist, ocplx* hess_vdirs)
{
	return orc_brickwall_hvp_columns(ufunc, udata, vlist, vdirs, nlayers, perms, L, 0, 1L << L, fval, dvlist, hess_vdirs);
}

/* reference src/target.c:631-678 */
int orc_brickwall_unitary_target_projected_hessian_vector_product(orc_linear_func ufunc, void* udata,
	const ocplx* vlist, const ocplx* vdirs, int nlayers, const int* perms, int L, ocplx* fval,
	double* grad_vec, double* hvp_vec)
{
	ocplx* dvlist = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	ocplx* hess = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	int ret = orc_brickwall_unitary_target_hessian_vector_product(ufunc, udata, vlist, vdirs, nlayers, perms, L, fval, dvlist, hess);
	if (ret < 0) { free(hess); free(dvlist); return ret; }
	orc_brickwall_project_grad_hvp(vlist, vdirs, nlayers, dvlist, hess, grad_vec, hvp_vec);
	free(hess); free(dvlist);
	return 0;
}

/* reference src/target.c:645-671 */
void orc_brickwall_project_grad_hvp(const ocplx* vlist, const ocplx* vdirs, int nlayers, ocplx* dvlist,
	ocplx* hess_vdirs, double* grad_vec, double* hvp_vec)
{
	for (int i = 0; i < nlayers; i++) {
		project_grad_hvp(&vlist[16*i], &vdirs[16*i], &dvlist[16*i], &hess_vdirs[16*i], &grad_vec[16*i], &hvp_vec[16*i]);
	}
}

/* ---------------------------------------------------------------------------
 * 4x4 algebra, reference src/matrix.c
 */

/* reference src/matrix.c:356-369 (i, j, k loop order) */
void orc_multiply(const ocplx* a, const ocplx* b, ocplx* c)
{
	for (int e = 0; e < 16; e++) { c[e] = 0; }
	for (int i = 0; i < 4; i++) {
		for (int j = 0; j < 4; j++) {
			for (int k = 0; k < 4; k++) {
				c[4*i + k] += a[4*i + j] * b[4*j + k];
			}
		}
	}
}

/* reference src/matrix.c:119-144 */
void orc_transpose(const ocplx* a, ocplx* at)
{
	for (int i = 0; i < 4; i++) { for (int j = 0; j < 4; j++) { at[4*i + j] = a[4*j + i]; } }
}
void orc_adjoint(const ocplx* a, ocplx* ah)
{
	for (int i = 0; i < 4; i++) { for (int j = 0; j < 4; j++) { ah[4*i + j] = conj(a[4*j + i]); } }
}

static const int upper_idx[6][2] = { {0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3} };

/* reference src/matrix.c:150-174 */
void orc_symm(const ocplx* w, ocplx* z)
{
	for (int k = 0; k < 4; k++) { z[5*k] = creal(w[5*k]); }
	for (int t = 0; t < 6; t++) {
		const int r = upper_idx[t][0], c = upper_idx[t][1];
		z[4*r + c] = 0.5 * (w[4*r + c] + conj(w[4*c + r]));
		z[4*c + r] = conj(z[4*r + c]);
	}
}

/* reference src/matrix.c:180-211 */
void orc_antisymm(const ocplx* w, ocplx* z)
{
	for (int k = 0; k < 4; k++) { z[5*k] = I * cimag(w[5*k]); }
	for (int t = 0; t < 6; t++) {
		const int r = upper_idx[t][0], c = upper_idx[t][1];
		z[4*r + c] = 0.5 * (w[4*r + c] - conj(w[4*c + r]));
		z[4*c + r] = -conj(z[4*r + c]);
	}
}

/* reference src/matrix.c:218-244 (complex case) */
void orc_real_to_antisymm(const double* r, ocplx* w)
{
	for (int k = 0; k < 4; k++) { w[5*k] = I * r[5*k]; }
	for (int t = 0; t < 6; t++) {
		const int a = upper_idx[t][0], b = upper_idx[t][1];
		const double p = r[4*a + b], q = r[4*b + a];
		w[4*a + b] = 0.5*(p - q) + 0.5*I*(p + q);
		w[4*b + a] = -conj(w[4*a + b]);
	}
}

/* reference src/matrix.c:279-288 (complex case) */
void orc_antisymm_to_real(const ocplx* w, double* r)
{
	for (int e = 0; e < 16; e++) { r[e] = creal(w[e]) + cimag(w[e]); }
}

/* reference src/matrix.c:309-314: z = v A(r) */
void orc_real_to_tangent(const double* r, const ocplx* v, ocplx* z)
{
	ocplx a[16];
	orc_real_to_antisymm(r, a);
	orc_multiply(v, a, z);
}

/* reference src/matrix.c:322-329: r = antisymm_to_real(antisymm(v^dagger z)) */
void orc_tangent_to_real(const ocplx* v, const ocplx* z, double* r)
{
	ocplx w[16], t[16];
	orc_adjoint(v, w);
	orc_multiply(w, z, t);
	orc_antisymm(t, w);
	orc_antisymm_to_real(w, r);
}

/* reference src/matrix.c:336-347: p = z - u symm(u^dagger z) */
void orc_project_tangent(const ocplx* u, const ocplx* z, ocplx* p)
{
	ocplx v[16], w[16];
	orc_adjoint(u, v);
	orc_multiply(v, z, w);
	orc_symm(w, v);
	orc_multiply(u, v, w);
	for (int e = 0; e < 16; e++) { p[e] = z[e] - w[e]; }
}

/* reference src/matrix.c:376-388: (abc + cba) / 2 */
void orc_symmetric_triple_matrix_product(const ocplx* a, const ocplx* b, const ocplx* c, ocplx* ret)
{
	ocplx u[16], v[16];
	orc_multiply(a, b, u);
	orc_multiply(u, c, ret);
	orc_multiply(c, b, u);
	orc_multiply(u, a, v);
	for (int e = 0; e < 16; e++) { ret[e] += v[e]; ret[e] *= 0.5; }
}

/* Gauss-Jordan elimination with partial pivoting, reference src/matrix.c:410-475 */
int orc_inverse_matrix(const ocplx* a, ocplx* ainv)
{
	ocplx m[16];
	memcpy(m, a, sizeof(m));
	for (int e = 0; e < 16; e++) { ainv[e] = (e % 5 == 0) ? 1 : 0; }
	for (int k = 0; k < 4; k++) {
		int piv = k;
		double p = cabs(m[5*k]);
		for (int i = k + 1; i < 4; i++) {
			if (cabs(m[4*i + k]) > p) { piv = i; p = cabs(m[4*i + k]); }
		}
		if (p == 0) { return -1; }
		if (piv != k) {
			for (int c = 0; c < 4; c++) {
				ocplx t = m[4*k + c]; m[4*k + c] = m[4*piv + c]; m[4*piv + c] = t;
				t = ainv[4*k + c]; ainv[4*k + c] = ainv[4*piv + c]; ainv[4*piv + c] = t;
			}
		}
		for (int i = 0; i < 4; i++) {
			if (i == k) { continue; }
			const ocplx s = m[4*i + k] / m[5*k];
			for (int c = k + 1; c < 4; c++) { m[4*i + c] -= s * m[4*k + c]; }
			m[4*i + k] = 0;
			for (int c = 0; c < 4; c++) { ainv[4*i + c] -= s * ainv[4*k + c]; }
		}
		for (int c = k + 1; c < 4; c++) { m[4*k + c] /= m[5*k]; }
		for (int c = 0; c < 4; c++) { ainv[4*k + c] /= m[5*k]; }
		m[5*k] = 1;
	}
	return 0;
}

/* Newton iteration u <- (u + u^{-dagger}) / 2 (Higham 1986), reference src/matrix.c:488-510 */
void orc_polar_factor(const ocplx* a, ocplx* u)
{
	memcpy(u, a, 16 * sizeof(ocplx));
	for (int k = 0; k < 14; k++) {
		ocplx v[16], w[16];
		orc_inverse_matrix(u, v);
		orc_adjoint(v, w);
		for (int e = 0; e < 16; e++) { u[e] += w[e]; u[e] *= 0.5; }
		if (k >= 4) {
			double d = 0;
			for (int e = 0; e < 16; e++) { d = fmax(d, cabs(u[e] - w[e])); }
			if (d < 1e-14) { break; }
		}
	}
}

/* reference src/matrix.c:517-531: v = polar(u (I + A(eta))) */
void orc_retract(const ocplx* u, const double* eta, ocplx* v)
{
	ocplx z[16], w[16];
	orc_real_to_antisymm(eta, z);
	for (int k = 0; k < 4; k++) { z[5*k] += 1; }
	orc_multiply(u, z, w);
	orc_polar_factor(w, v);
}

/* ---------------------------------------------------------------------------
 * Trust region, reference src/trust_region.c (sequential dot products in
 * place of CBLAS ddot/daxpy/dscal).
 */
static double ddot(int n, const double* x, const double* y)
{
	double s = 0;
	for (int i = 0; i < n; i++) { s += x[i] * y[i]; }
	return s;
}

/* roots of x^2 + 2 p x + q = 0 in ascending order, reference src/trust_region.c:14-40 */
static void solve_quadratic(double p, double q, double x[2])
{
	if (p == 0) {
		const double a = sqrt(-q);
		x[0] = -a; x[1] = a;
		return;
	}
	const double a = -(p + copysign(1.0, p) * sqrt(p * p - q));
	const double b = q / a;
	if (a <= b) { x[0] = a; x[1] = b; } else { x[0] = b; x[1] = a; }
}

/* t > 0 with ||b + t d|| = radius, reference src/trust_region.c:49-65 */
static double move_to_boundary(const double* b, const double* d, int n, double radius)
{
	const double dsq = ddot(n, d, d);
	if (dsq == 0) { return 0; }
	const double p = ddot(n, b, d) / dsq;
	const double q = (ddot(n, b, b) - radius * radius) / dsq;
	double t[2];
	solve_quadratic(p, q, t);
	return t[1];
}

/* Steihaug-Toint truncated CG, reference src/trust_region.c:159-217 */
bool orc_truncated_cg_hvp(const double* x, const double* grad, orc_hvp_func f_hvp, void* fdata, int n,
	double radius, int maxiter, double abstol, double reltol, double* z)
{
	double* r = xalloc(n * sizeof(double));
	double* d = xalloc(n * sizeof(double));
	double* hd = xalloc(n * sizeof(double));
	memcpy(r, grad, n * sizeof(double));
	double rsq = ddot(n, r, r);
	const double stoptol = fmax(abstol, reltol * sqrt(rsq));
	for (int i = 0; i < n; i++) { z[i] = 0; d[i] = -r[i]; }
	bool boundary = false;
	for (int it = 0; it < maxiter; it++) {
		f_hvp(x, fdata, d, hd);
		const double dhd = ddot(n, d, hd);
		const double t = move_to_boundary(z, d, n, radius);
		const double alpha = rsq / dhd;
		if (dhd <= 0 || alpha > t) {
			for (int i = 0; i < n; i++) { z[i] += t * d[i]; }
			boundary = true;
			break;
		}
		for (int i = 0; i < n; i++) { r[i] += alpha * hd[i]; }
		for (int i = 0; i < n; i++) { z[i] += alpha * d[i]; }
		const double rsq_next = ddot(n, r, r);
		if (sqrt(rsq_next) <= stoptol) { break; }
		const double beta = rsq_next / rsq;
		for (int i = 0; i < n; i++) { d[i] = beta * d[i] - r[i]; }
		rsq = rsq_next;
	}
	free(hd); free(d); free(r);
	return boundary;
}

struct bw_data { orc_linear_func ufunc; void* udata; const int* perms; int nlayers; int L; };

/* wrappers, reference src/brickwall_opt.c:22-79 */
static double bw_f(const double* x, struct bw_data* dd)
{
	ocplx f;
	if (orc_brickwall_unitary_target(dd->ufunc, dd->udata, (const ocplx*)x, dd->nlayers, dd->L, dd->perms, &f) < 0) { return -1; }
	return creal(f);
}
static double bw_f_deriv(const double* x, struct bw_data* dd, double* grad)
{
	ocplx f;
	if (orc_brickwall_unitary_target_and_projected_gradient(dd->ufunc, dd->udata, (const ocplx*)x, dd->nlayers, dd->L, dd->perms, &f, grad) < 0) { return -1; }
	return creal(f);
}
static void bw_f_hvp(const double* x, void* fdata, const double* vec, double* hvp)
{
	struct bw_data* dd = fdata;
	const ocplx* vlist = (const ocplx*)x;
	ocplx* vdirs = xalloc((size_t)dd->nlayers * 16 * sizeof(ocplx));
	double* grad = xalloc((size_t)dd->nlayers * 16 * sizeof(double));
	for (int i = 0; i < dd->nlayers; i++) { orc_real_to_tangent(&vec[16*i], &vlist[16*i], &vdirs[16*i]); }
	ocplx f;
	orc_brickwall_unitary_target_projected_hessian_vector_product(dd->ufunc, dd->udata, vlist, vdirs, dd->nlayers, dd->perms, dd->L, &f, grad, hvp);
	free(grad); free(vdirs);
}

/* Riemannian trust region (Absil et al. Alg. 10) with default parameters,
   reference src/trust_region.c:306-368 and src/brickwall_opt.c:107-150 */
void orc_optimize_brickwall_circuit_hvp(orc_linear_func ufunc, void* udata, const ocplx* vlist_start, int nlayers,
	int L, const int* perms, int niter, double* f_iter, ocplx* vlist_opt)
{
	struct bw_data dd = { ufunc, udata, perms, nlayers, L };
	const int n = 16 * nlayers, m = 32 * nlayers;
	const int maxiter = 2 * n;
	const double abstol = 1e-8, reltol = 1e-6, rho_trust = 0.125, maxradius = 0.1;
	double radius = 0.01;
	double* x = xalloc(m * sizeof(double));
	double* xn = xalloc(m * sizeof(double));
	double* grad = xalloc(n * sizeof(double));
	double* eta = xalloc(n * sizeof(double));
	double* heta = xalloc(n * sizeof(double));
	memcpy(x, vlist_start, m * sizeof(double));
	for (int k = 0; k < niter; k++) {
		const double fx = bw_f_deriv(x, &dd, grad);
		const bool on_boundary = orc_truncated_cg_hvp(x, grad, bw_f_hvp, &dd, n, radius, maxiter, abstol, reltol, eta);
		for (int i = 0; i < nlayers; i++) { orc_retract((const ocplx*)&x[32*i], &eta[16*i], (ocplx*)&xn[32*i]); }
		f_iter[k] = fx;
		bw_f_hvp(x, &dd, eta, heta);
		const double rho = (bw_f(xn, &dd) - fx) / (ddot(n, grad, eta) + 0.5 * ddot(n, eta, heta));
		if (rho < 0.25) { radius *= 0.25; }
		else if (rho > 0.75 && on_boundary) { radius = fmin(2 * radius, maxradius); }
		if (rho > rho_trust) { double* t = x; x = xn; xn = t; }
	}
	memcpy(vlist_opt, x, m * sizeof(double));
	f_iter[niter] = bw_f(x, &dd);
	free(heta); free(eta); free(grad); free(xn); free(x);
}

struct qc_data { orc_linear_func ufunc; void* udata; const int* wires; int ngates; int L; };

/* wrappers, reference src/circuit_opt.c:22-79 */
static double qc_f(const double* x, struct qc_data* dd)
{
	ocplx f;
	if (orc_circuit_unitary_target(dd->ufunc, dd->udata, (const ocplx*)x, dd->ngates, dd->wires, dd->L, &f) < 0) { return -1; }
	return creal(f);
}
static double qc_f_deriv(const double* x, struct qc_data* dd, double* grad)
{
	ocplx f;
	if (orc_circuit_unitary_target_and_projected_gradient(dd->ufunc, dd->udata, (const ocplx*)x, dd->ngates, dd->wires, dd->L, &f, grad) < 0) { return -1; }
	return creal(f);
}
static void qc_f_hvp(const double* x, void* fdata, const double* vec, double* hvp)
{
	struct qc_data* dd = fdata;
	const ocplx* gates = (const ocplx*)x;
	ocplx* dirs = xalloc((size_t)dd->ngates * 16 * sizeof(ocplx));
	double* grad = xalloc((size_t)dd->ngates * 16 * sizeof(double));
	for (int i = 0; i < dd->ngates; i++) { orc_real_to_tangent(&vec[16*i], &gates[16*i], &dirs[16*i]); }
	ocplx f;
	orc_circuit_unitary_target_projected_hessian_vector_product(dd->ufunc, dd->udata, gates, dirs, dd->ngates, dd->wires, dd->L, &f, grad, hvp);
	free(grad); free(dirs);
}

/* reference src/circuit_opt.c:104-124 with src/trust_region.c:306-368, default parameters */
void orc_optimize_quantum_circuit(orc_linear_func ufunc, void* udata, const ocplx* gates_start, int ngates, int L,
	const int* wires, int niter, double* f_iter, ocplx* gates_opt)
{
	struct qc_data dd = { ufunc, udata, wires, ngates, L };
	const int n = 16 * ngates, m = 32 * ngates;
	const int maxiter = 2 * n;
	const double abstol = 1e-8, reltol = 1e-6, rho_trust = 0.125, maxradius = 0.1;
	double radius = 0.01;
	double* x = xalloc(m * sizeof(double));
	double* xn = xalloc(m * sizeof(double));
	double* grad = xalloc(n * sizeof(double));
	double* eta = xalloc(n * sizeof(double));
	double* heta = xalloc(n * sizeof(double));
	memcpy(x, gates_start, m * sizeof(double));
	for (int k = 0; k < niter; k++) {
		const double fx = qc_f_deriv(x, &dd, grad);
		const bool on_boundary = orc_truncated_cg_hvp(x, grad, qc_f_hvp, &dd, n, radius, maxiter, abstol, reltol, eta);
		for (int i = 0; i < ngates; i++) { orc_retract((const ocplx*)&x[32*i], &eta[16*i], (ocplx*)&xn[32*i]); }
		f_iter[k] = fx;
		qc_f_hvp(x, &dd, eta, heta);
		const double rho = (qc_f(xn, &dd) - fx) / (ddot(n, grad, eta) + 0.5 * ddot(n, eta, heta));
		if (rho < 0.25) { radius *= 0.25; }
		else if (rho > 0.75 && on_boundary) { radius = fmin(2 * radius, maxradius); }
		if (rho > rho_trust) { double* t = x; x = xn; xn = t; }
	}
	memcpy(gates_opt, x, m * sizeof(double));
	f_iter[niter] = qc_f(x, &dd);
	free(heta); free(eta); free(grad); free(xn); free(x);
}

/* ---------------------------------------------------------------------------
 * Full-Hessian route.
 */

/* the four amplitude offsets of the quad whose other bits are those of `base`,
   in the gate basis |q_i q_j> (wire i the more significant index) */
static void quad_offsets4(long base, int L, int i, int j, long off[4])
{
	const long si = 1L << (L - 1 - i), sj = 1L << (L - 1 - j);
	off[0] = base; off[1] = base + sj; off[2] = base + si; off[3] = base + si + sj;
}

/* reference src/gate.c:374-460: state 4r + c of the output array is E_rc psi,
   i.e. psi's (i, j) = c component moved to (i, j) = r */
void orc_apply_gate_placeholder(int i, int j, int L, const ocplx* psi, ocplx* psi_out)
{
	const long n = 1L << L;
	memset(psi_out, 0, (size_t)n * 16 * sizeof(ocplx));
	const int lo = i < j ? i : j, hi = i < j ? j : i;
	const long nq = 1L << (L - 2);
	for (long q = 0; q < nq; q++) {
		long off[4];
		quad_offsets4(quad_base(q, L, lo, hi), L, i, j, off);
		for (int r = 0; r < 4; r++) {
			for (int c = 0; c < 4; c++) {
				psi_out[off[r] * 16 + 4*r + c] = psi[off[c]];
			}
		}
	}
}

/* reference src/gate.c:269-366 */
void orc_apply_gate_backward_array(int i, int j, int L, int nstates, const ocplx* psi, const ocplx* dpsi_out, ocplx* dgates)
{
	for (int e = 0; e < 16 * nstates; e++) { dgates[e] = 0; }
	const int lo = i < j ? i : j, hi = i < j ? j : i;
	const long nq = 1L << (L - 2);
	for (long q = 0; q < nq; q++) {
		long off[4];
		quad_offsets4(quad_base(q, L, lo, hi), L, i, j, off);
		for (int k = 0; k < nstates; k++) {
			for (int r = 0; r < 4; r++) {
				for (int c = 0; c < 4; c++) {
					dgates[16*k + 4*r + c] += dpsi_out[off[r]] * psi[off[c] * nstates + k];
				}
			}
		}
	}
}

/* reference src/brickwall_circuit.c:134-255.  Gradient sweep storing the adjoint
   after every gate; then, for every gate position, the 16 placeholder states are
   carried forward through the later gates and contracted with the adjoint at
   each of them (pairs of one layer: only later pairs); diagonal blocks are
   symmetrised and the lower blocks copied from the upper ones. */
int orc_brickwall_unitary_backward_hessian(const ocplx* vlist, int nlayers, const int* perms, int L, const ocplx* cache,
	const ocplx* dpsi_out, ocplx* dpsi, ocplx* dvlist, ocplx* hess)
{
	const long n = 1L << L;
	const int per = L / 2, ng = nlayers * per, m = 16 * nlayers;
	ocplx* adj = xalloc((size_t)ng * n * sizeof(ocplx));   /* adjoint after gate g */
	memcpy(&adj[(size_t)(ng - 1) * n], dpsi_out, n * sizeof(ocplx));
	for (int i = 0; i < nlayers; i++) { for (int e = 0; e < 16; e++) { dvlist[16*i + e] = 0; } }
	for (int g = ng - 1; g >= 0; g--) {
		const int i = g / per, k = g % per;
		ocplx dV[16];
		ocplx* out = (g > 0) ? &adj[(size_t)(g - 1) * n] : dpsi;
		orc_apply_gate_backward(&vlist[16*i], perms[i*L + 2*k], perms[i*L + 2*k + 1], L, &cache[(size_t)g * n],
			&adj[(size_t)g * n], out, dV);
		for (int e = 0; e < 16; e++) { dvlist[16*i + e] += dV[e]; }
	}
	memset(hess, 0, (size_t)m * m * sizeof(ocplx));
	ocplx* t0 = xalloc((size_t)n * 16 * sizeof(ocplx));
	ocplx* t1 = xalloc((size_t)n * 16 * sizeof(ocplx));
	ocplx h[256];
	for (int g = 0; g < ng; g++) {
		const int i = g / per, r = g % per;
		orc_apply_gate_placeholder(perms[i*L + 2*r], perms[i*L + 2*r + 1], L, &cache[(size_t)g * n], t0);
		ocplx* cur = t0;
		ocplx* nxt = t1;
		for (int l = g + 1; l < ng; l++) {
			const int j = l / per, s2 = l % per;
			const int wi = perms[j*L + 2*s2], wj = perms[j*L + 2*s2 + 1];
			orc_apply_gate_backward_array(wi, wj, L, 16, cur, &adj[(size_t)l * n], h);
			for (int x = 0; x < 16; x++) {
				for (int y = 0; y < 16; y++) {
					hess[(size_t)(i*16 + x) * m + j*16 + y] += h[16*x + y];
				}
			}
			if (l < ng - 1) {
				orc_apply_gate_to_array(&vlist[16*j], wi, wj, L, 16, cur, nxt);
				ocplx* t = cur; cur = nxt; nxt = t;
			}
		}
	}
	for (int i = 0; i < nlayers; i++) {
		for (int x = 0; x < 16; x++) {
			for (int y = 0; y <= x; y++) {
				const size_t a = (size_t)(i*16 + x) * m + i*16 + y, b = (size_t)(i*16 + y) * m + i*16 + x;
				const ocplx sum = hess[a] + hess[b];
				hess[a] = sum;
				hess[b] = sum;
			}
		}
		for (int j = 0; j < i; j++) {
			for (int x = 0; x < 16; x++) {
				for (int y = 0; y < 16; y++) {
					hess[(size_t)(i*16 + x) * m + j*16 + y] = hess[(size_t)(j*16 + y) * m + i*16 + x];
				}
			}
		}
	}
	free(t1); free(t0); free(adj);
	return 0;
}

/* reference src/target.c:687-797: per basis column, forward with cache, trace
   against -conj(U e_b), backward Hessian; all summed over the columns */
int orc_brickwall_unitary_target_gradient_hessian(orc_linear_func ufunc, void* udata, const ocplx* vlist, int nlayers,
	int L, const int* perms, ocplx* fval, ocplx* dvlist, ocplx* hess)
{
	const long n = 1L << L;
	const int ng = nlayers * (L / 2), m = 16 * nlayers;
	ocplx* psi = xalloc(n * sizeof(ocplx));
	ocplx* upsi = xalloc(n * sizeof(ocplx));
	ocplx* wpsi = xalloc(n * sizeof(ocplx));
	ocplx* cache = xalloc((size_t)ng * n * sizeof(ocplx));
	ocplx* dvu = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	ocplx* hu = xalloc((size_t)m * m * sizeof(ocplx));
	for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] = 0; }
	memset(hess, 0, (size_t)m * m * sizeof(ocplx));
	ocplx f = 0;
	int ret = 0;
	for (long b = 0; b < n && ret == 0; b++) {
		if (adjoint_seed(ufunc, udata, L, b, psi, upsi) < 0) { ret = -2; break; }
		orc_brickwall_unitary_forward(vlist, nlayers, perms, L, psi, cache, wpsi);
		for (long a = 0; a < n; a++) { f += upsi[a] * wpsi[a]; }
		orc_brickwall_unitary_backward_hessian(vlist, nlayers, perms, L, cache, upsi, psi, dvu, hu);
		for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] += dvu[e]; }
		for (size_t e = 0; e < (size_t)m * m; e++) { hess[e] += hu[e]; }
	}
	*fval = f;
	free(hu); free(dvu); free(cache); free(wpsi); free(upsi); free(psi);
	return ret;
}

/* reference src/target.c:937-994: upper blocks (i <= j) of the Riemannian Hessian,
   column k of block (i, j) from the tangent basis vector k at V_j; the diagonal
   blocks also carry the curvature term of the gradient; lower half by symmetry.
   hess2 (NULL for the unitary targets) is the Hessian of the conjugated gates,
   added unconjugated as in reference src/target.c:1665-1680. */
static void project_hessian_matrix(int nlayers, const ocplx* vlist, const ocplx* dvc, const ocplx* hess,
	const ocplx* hess2, double* H)
{
	const int m = 16 * nlayers;
	for (int i = 0; i < nlayers; i++) {
		for (int j = i; j < nlayers; j++) {
			for (int k = 0; k < 16; k++) {
				double r[16] = { 0 };
				r[k] = 1;
				ocplx Z[16], G[16];
				orc_real_to_tangent(r, &vlist[16*j], Z);
				for (int x = 0; x < 16; x++) {
					ocplx s = 0;
					for (int y = 0; y < 16; y++) { s += hess[(size_t)(i*16 + x) * m + j*16 + y] * Z[y]; }
					G[x] = conj(s);
					if (hess2 != NULL) {
						ocplx s2 = 0;
						for (int y = 0; y < 16; y++) { s2 += hess2[(size_t)(i*16 + x) * m + j*16 + y] * Z[y]; }
						G[x] += s2;
					}
				}
				if (i == j) {
					ocplx P[16], gh[16], T[16];
					orc_project_tangent(&vlist[16*i], G, P);
					orc_adjoint(&dvc[16*i], gh);
					orc_symmetric_triple_matrix_product(Z, gh, &vlist[16*i], T);
					for (int e = 0; e < 16; e++) { G[e] = P[e] - T[e]; }
				}
				orc_tangent_to_real(&vlist[16*i], G, r);
				for (int x = 0; x < 16; x++) { H[(size_t)(i*16 + x) * m + j*16 + k] = r[x]; }
			}
		}
	}
	for (int a = 0; a < m; a++) {
		for (int b = 0; b < a; b++) { H[(size_t)a * m + b] = H[(size_t)b * m + a]; }
	}
}

/* reference src/target.c:1003-1040 */
int orc_brickwall_unitary_target_gradient_vector_hessian_matrix(orc_linear_func ufunc, void* udata, const ocplx* vlist,
	int nlayers, int L, const int* perms, ocplx* fval, double* grad_vec, double* H)
{
	const int m = 16 * nlayers;
	ocplx* dv = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	ocplx* hess = xalloc((size_t)m * m * sizeof(ocplx));
	int ret = orc_brickwall_unitary_target_gradient_hessian(ufunc, udata, vlist, nlayers, L, perms, fval, dv, hess);
	if (ret == 0) {
		for (int i = 0; i < nlayers; i++) {
			for (int e = 0; e < 16; e++) { dv[16*i + e] = conj(dv[16*i + e]); }
			orc_tangent_to_real(&vlist[16*i], &dv[16*i], &grad_vec[16*i]);
		}
		project_hessian_matrix(nlayers, vlist, dv, hess, NULL, H);
	}
	free(hess); free(dv);
	return ret;
}

/* reference src/trust_region.c:83-141 (Steihaug tCG with an explicit Hessian) */
bool orc_truncated_cg_hmat(const double* grad, const double* hess, int n, double radius, int maxiter, double abstol,
	double reltol, double* z)
{
	double* r = xalloc(n * sizeof(double));
	double* d = xalloc(n * sizeof(double));
	double* hd = xalloc(n * sizeof(double));
	memcpy(r, grad, n * sizeof(double));
	double rsq = ddot(n, r, r);
	const double stoptol = fmax(abstol, reltol * sqrt(rsq));
	for (int i = 0; i < n; i++) { z[i] = 0; d[i] = -r[i]; }
	bool boundary = false;
	for (int it = 0; it < maxiter; it++) {
		for (int a = 0; a < n; a++) { hd[a] = ddot(n, &hess[(size_t)a * n], d); }
		const double dhd = ddot(n, d, hd);
		const double t = move_to_boundary(z, d, n, radius);
		const double alpha = rsq / dhd;
		if (dhd <= 0 || alpha > t) {
			for (int i = 0; i < n; i++) { z[i] += t * d[i]; }
			boundary = true;
			break;
		}
		for (int i = 0; i < n; i++) { r[i] += alpha * hd[i]; }
		for (int i = 0; i < n; i++) { z[i] += alpha * d[i]; }
		const double rsq_next = ddot(n, r, r);
		if (sqrt(rsq_next) <= stoptol) { break; }
		const double beta = rsq_next / rsq;
		for (int i = 0; i < n; i++) { d[i] = beta * d[i] - r[i]; }
		rsq = rsq_next;
	}
	free(hd); free(d); free(r);
	return boundary;
}

/* RTR with the full Hessian (reference src/trust_region.c:230-293) on the
   brickwall target (reference src/brickwall_opt.c f_deriv_hess and
   optimize_brickwall_circuit_hmat), default parameters */
void orc_optimize_brickwall_circuit_hmat(orc_linear_func ufunc, void* udata, const ocplx* vlist_start, int nlayers,
	int L, const int* perms, int niter, double* f_iter, ocplx* vlist_opt)
{
	struct bw_data dd = { ufunc, udata, perms, nlayers, L };
	const int n = 16 * nlayers, m = 32 * nlayers;
	const int maxiter = 2 * n;
	const double abstol = 1e-8, reltol = 1e-6, rho_trust = 0.125, maxradius = 0.1;
	double radius = 0.01;
	double* x = xalloc(m * sizeof(double));
	double* xn = xalloc(m * sizeof(double));
	double* grad = xalloc(n * sizeof(double));
	double* H = xalloc((size_t)n * n * sizeof(double));
	double* eta = xalloc(n * sizeof(double));
	double* heta = xalloc(n * sizeof(double));
	memcpy(x, vlist_start, m * sizeof(double));
	for (int k = 0; k < niter; k++) {
		ocplx fv;
		orc_brickwall_unitary_target_gradient_vector_hessian_matrix(ufunc, udata, (const ocplx*)x, nlayers, L, perms, &fv, grad, H);
		const double fx = creal(fv);
		const bool on_boundary = orc_truncated_cg_hmat(grad, H, n, radius, maxiter, abstol, reltol, eta);
		for (int i = 0; i < nlayers; i++) { orc_retract((const ocplx*)&x[32*i], &eta[16*i], (ocplx*)&xn[32*i]); }
		f_iter[k] = fx;
		for (int a = 0; a < n; a++) { heta[a] = ddot(n, &H[(size_t)a * n], eta); }
		const double rho = (bw_f(xn, &dd) - fx) / (ddot(n, grad, eta) + 0.5 * ddot(n, eta, heta));
		if (rho < 0.25) { radius *= 0.25; }
		else if (rho > 0.75 && on_boundary) { radius = fmin(2 * radius, maxradius); }
		if (rho > rho_trust) { double* t = x; x = xn; xn = t; }
	}
	memcpy(vlist_opt, x, m * sizeof(double));
	f_iter[niter] = bw_f(x, &dd);
	free(heta); free(eta); free(H); free(grad); free(xn); free(x);
}

/* ---------------------------------------------------------------------------
 * RNG: PCG32 (O'Neill 2014, XSH-RR output), two generators combined to 64 bits,
 * reference src/pcg_basic.c:30-51,112-132 and src/rng.c:13-25,91-123,152-157.
 */
static uint32_t pcg32_next(struct orc_pcg32* g)
{
	const uint64_t old = g->state;
	g->state = old * 6364136223846793005ULL + g->inc;
	const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
	const uint32_t rot = (uint32_t)(old >> 59u);
	return (xs >> rot) | (xs << ((-rot) & 31));
}

static void pcg32_seed(struct orc_pcg32* g, uint64_t initstate, uint64_t initseq)
{
	g->state = 0;
	g->inc = (initseq << 1u) | 1u;
	pcg32_next(g);
	g->state += initstate;
	pcg32_next(g);
}

static uint64_t pcg32_pair(struct orc_pcg32* g)
{
	const uint64_t hi = pcg32_next(g);
	const uint64_t lo = pcg32_next(g);
	return (hi << 32) | lo;
}

void orc_seed_rng_state(uint64_t seed, struct orc_rng* rng)
{
	struct orc_pcg32 tmp;
	pcg32_seed(&tmp, seed, 6364136223846793005ULL * seed + 1442695040888963407ULL);
	const uint64_t s1 = pcg32_pair(&tmp);
	const uint64_t s2 = pcg32_pair(&tmp);
	const uint64_t q1 = pcg32_pair(&tmp);
	uint64_t q2 = pcg32_pair(&tmp);
	const uint64_t mask = ~0ULL >> 1;
	if ((q1 & mask) == (q2 & mask)) { q2 = ~q2; }
	pcg32_seed(&rng->gen[0], s1, q1);
	pcg32_seed(&rng->gen[1], s2, q2);
}

uint32_t orc_rand_uint32(struct orc_rng* rng) { return pcg32_next(&rng->gen[0]); }

uint64_t orc_rand_uint64(struct orc_rng* rng)
{
	const uint64_t hi = pcg32_next(&rng->gen[0]);
	const uint64_t lo = pcg32_next(&rng->gen[1]);
	return (hi << 32) | lo;
}

double orc_randu(struct orc_rng* rng) { return ldexp((double)orc_rand_uint64(rng), -64); }

/* Box-Muller (1958) */
double orc_randn(struct orc_rng* rng)
{
	double u1 = orc_randu(rng);
	const double u2 = orc_randu(rng);
	if (u1 == 0) { u1 = 1; }
	return sqrt(-2 * log(u1)) * sin(6.2831853071795864769 * u2);
}

ocplx orc_crandn(struct orc_rng* rng)
{
	const double x = orc_randn(rng);
	const double y = orc_randn(rng);
	return (x + I * y) / sqrt(2.0);
}

/* ---------------------------------------------------------------------------
 * Statevector utilities and sampled (stochastic-trace) targets.
 */

/* reference src/statevector.c:47-66 (Euclidean norm, then scaling by its inverse) */
void orc_normalize_statevector(int L, ocplx* psi)
{
	const long n = 1L << L;
	double s = 0;
	for (long j = 0; j < n; j++) { s += creal(psi[j]) * creal(psi[j]) + cimag(psi[j]) * cimag(psi[j]); }
	const double nrm = sqrt(s);
	if (nrm > 0) {
		for (long j = 0; j < n; j++) { psi[j] *= 1.0 / nrm; }
	}
}

/* reference src/statevector.c:72-92: output bit i of the index reads input bit L-1-perm[L-1-i] */
void orc_transpose_statevector(int L, const ocplx* psi, const int* perm, ocplx* psi_trans)
{
	const long n = 1L << L;
	for (long j = 0; j < n; j++) {
		long k = 0;
		for (int i = 0; i < L; i++) {
			if ((j >> i) & 1) { k |= 1L << (L - 1 - perm[L - 1 - i]); }
		}
		psi_trans[j] = psi[k];
	}
}

/* reference src/statevector.c:102-127: crandn entries in order, then normalised */
void orc_haar_random_statevector(int L, struct orc_rng* rng, ocplx* psi)
{
	const long n = 1L << L;
	for (long j = 0; j < n; j++) { psi[j] = orc_crandn(rng); }
	orc_normalize_statevector(L, psi);
}

/* reference src/target.c:85-149: sum over samples of <-conj(U psi), C psi>, scaled by 2^L / nsamples */
int orc_circuit_unitary_target_sampling(orc_linear_func ufunc, void* udata, const ocplx* gates, int ngates, const int* wires,
	int L, long nsamples, struct orc_rng* rng, ocplx* fval)
{
	const long n = 1L << L;
	ocplx* psi = xalloc(n * sizeof(ocplx));
	ocplx* upsi = xalloc(n * sizeof(ocplx));
	ocplx* cpsi = xalloc(n * sizeof(ocplx));
	ocplx f = 0;
	int ret = 0;
	for (long b = 0; b < nsamples; b++) {
		orc_haar_random_statevector(L, rng, psi);
		if (ufunc(L, psi, udata, upsi) < 0) { ret = -2; break; }
		orc_apply_quantum_circuit(gates, ngates, wires, L, psi, cpsi);
		for (long a = 0; a < n; a++) { f += -conj(upsi[a]) * cpsi[a]; }
	}
	*fval = f * ((double)n / nsamples);
	free(cpsi); free(upsi); free(psi);
	return ret;
}

/* reference src/target.c:491-505 */
int orc_brickwall_unitary_target_sampling(orc_linear_func ufunc, void* udata, const ocplx* vlist, int nlayers, int L,
	const int* perms, long nsamples, struct orc_rng* rng, ocplx* fval)
{
	const int ng = nlayers * (L / 2);
	ocplx* gates = xalloc((size_t)ng * 16 * sizeof(ocplx));
	int* wires = xalloc((size_t)ng * 2 * sizeof(int));
	brickwall_to_sequential(L, nlayers, vlist, perms, gates, wires);
	int ret = orc_circuit_unitary_target_sampling(ufunc, udata, gates, ng, wires, L, nsamples, rng, fval);
	free(wires); free(gates);
	return ret;
}

/* reference src/target.c:807-928: per sample, forward with cache, trace, backward Hessian; scaled sums */
int orc_brickwall_unitary_target_gradient_hessian_sampling(orc_linear_func ufunc, void* udata, const ocplx* vlist,
	int nlayers, int L, const int* perms, long nsamples, struct orc_rng* rng, ocplx* fval, ocplx* dvlist, ocplx* hess)
{
	const long n = 1L << L;
	const int ng = nlayers * (L / 2), m = 16 * nlayers;
	ocplx* psi = xalloc(n * sizeof(ocplx));
	ocplx* upsi = xalloc(n * sizeof(ocplx));
	ocplx* wpsi = xalloc(n * sizeof(ocplx));
	ocplx* cache = xalloc((size_t)ng * n * sizeof(ocplx));
	ocplx* dvu = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	ocplx* hu = xalloc((size_t)m * m * sizeof(ocplx));
	for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] = 0; }
	memset(hess, 0, (size_t)m * m * sizeof(ocplx));
	ocplx f = 0;
	int ret = 0;
	for (long b = 0; b < nsamples; b++) {
		orc_haar_random_statevector(L, rng, psi);
		if (ufunc(L, psi, udata, upsi) < 0) { ret = -2; break; }
		for (long a = 0; a < n; a++) { upsi[a] = -conj(upsi[a]); }
		orc_brickwall_unitary_forward(vlist, nlayers, perms, L, psi, cache, wpsi);
		for (long a = 0; a < n; a++) { f += upsi[a] * wpsi[a]; }
		orc_brickwall_unitary_backward_hessian(vlist, nlayers, perms, L, cache, upsi, psi, dvu, hu);
		for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] += dvu[e]; }
		for (size_t e = 0; e < (size_t)m * m; e++) { hess[e] += hu[e]; }
	}
	const double scale = (double)n / nsamples;
	*fval = f * scale;
	for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] *= scale; }
	for (size_t e = 0; e < (size_t)m * m; e++) { hess[e] *= scale; }
	free(hu); free(dvu); free(cache); free(wpsi); free(upsi); free(psi);
	return ret;
}

/* reference src/target.c:1050-1088 */
int orc_brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(orc_linear_func ufunc, void* udata,
	const ocplx* vlist, int nlayers, int L, const int* perms, long nsamples, struct orc_rng* rng, ocplx* fval,
	double* grad_vec, double* H)
{
	const int m = 16 * nlayers;
	ocplx* dv = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	ocplx* hess = xalloc((size_t)m * m * sizeof(ocplx));
	int ret = orc_brickwall_unitary_target_gradient_hessian_sampling(ufunc, udata, vlist, nlayers, L, perms, nsamples, rng,
		fval, dv, hess);
	if (ret == 0) {
		for (int i = 0; i < nlayers; i++) {
			for (int e = 0; e < 16; e++) { dv[16*i + e] = conj(dv[16*i + e]); }
			orc_tangent_to_real(&vlist[16*i], &dv[16*i], &grad_vec[16*i]);
		}
		project_hessian_matrix(nlayers, vlist, dv, hess, NULL, H);
	}
	free(hess); free(dv);
	return ret;
}

/* reference src/brickwall_opt.c:176-268 with src/trust_region.c:230-293: every target and
   derivative evaluation draws fresh samples from rng; default parameters */
void orc_optimize_brickwall_circuit_hmat_sampling(orc_linear_func ufunc, void* udata, const ocplx* vlist_start,
	int nlayers, int L, const int* perms, long nsamples, struct orc_rng* rng, int niter, double* f_iter, ocplx* vlist_opt)
{
	const int n = 16 * nlayers, m = 32 * nlayers;
	const int maxiter = 2 * n;
	const double abstol = 1e-8, reltol = 1e-6, rho_trust = 0.125, maxradius = 0.1;
	double radius = 0.01;
	double* x = xalloc(m * sizeof(double));
	double* xn = xalloc(m * sizeof(double));
	double* grad = xalloc(n * sizeof(double));
	double* H = xalloc((size_t)n * n * sizeof(double));
	double* eta = xalloc(n * sizeof(double));
	double* heta = xalloc(n * sizeof(double));
	memcpy(x, vlist_start, m * sizeof(double));
	for (int k = 0; k < niter; k++) {
		ocplx fv;
		orc_brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(ufunc, udata, (const ocplx*)x, nlayers, L, perms,
			nsamples, rng, &fv, grad, H);
		const double fx = creal(fv);
		const bool on_boundary = orc_truncated_cg_hmat(grad, H, n, radius, maxiter, abstol, reltol, eta);
		for (int i = 0; i < nlayers; i++) { orc_retract((const ocplx*)&x[32*i], &eta[16*i], (ocplx*)&xn[32*i]); }
		f_iter[k] = fx;
		for (int a = 0; a < n; a++) { heta[a] = ddot(n, &H[(size_t)a * n], eta); }
		ocplx fn;
		orc_brickwall_unitary_target_sampling(ufunc, udata, (const ocplx*)xn, nlayers, L, perms, nsamples, rng, &fn);
		const double rho = (creal(fn) - fx) / (ddot(n, grad, eta) + 0.5 * ddot(n, eta, heta));
		if (rho < 0.25) { radius *= 0.25; }
		else if (rho > 0.75 && on_boundary) { radius = fmin(2 * radius, maxradius); }
		if (rho > rho_trust) { double* t = x; x = xn; xn = t; }
	}
	memcpy(vlist_opt, x, m * sizeof(double));
	ocplx ff;
	orc_brickwall_unitary_target_sampling(ufunc, udata, (const ocplx*)x, nlayers, L, perms, nsamples, rng, &ff);
	f_iter[niter] = creal(ff);
	free(heta); free(eta); free(H); free(grad); free(xn); free(x);
}

/* ---------------------------------------------------------------------------
 * Block-encoding targets f = || P^dagger W P - H ||_F^2 / 2 (reference
 * src/target.c:1100-1720).  P embeds an L/2-qubit state on the odd bits of the
 * amplitude index (bit i of b goes to bit 2i + 1), the even bits held at |0>;
 * the Frobenius norm is summed over the 2^(L/2) unit vectors e_b.
 */

/* reference src/target.c:1146-1155: c = sum_i 2 t_i^2 with t_i = b & 2^i */
static long blockenc_embed(long b, int half)
{
	long c = 0;
	for (int i = 0; i < half; i++) {
		const long t = b & (1L << i);
		c += 2 * t * t;
	}
	return c;
}

/* H e_b into hpsi (2^(L/2) entries) and the embedded unit vector e_{c(b)} into chi */
static int blockenc_column(orc_linear_func hfunc, void* hdata, int L, long b, ocplx* psi, ocplx* hpsi, ocplx* chi)
{
	const int half = L / 2;
	const long s = 1L << half, n = 1L << L;
	memset(psi, 0, s * sizeof(ocplx));
	psi[b] = 1;
	if (hfunc(half, psi, hdata, hpsi) < 0) { return -2; }
	memset(chi, 0, n * sizeof(ocplx));
	chi[blockenc_embed(b, half)] = 1;
	return 0;
}

/* residual d_a = (W P e_b)[c(a)] - (H e_b)[a] (reference src/target.c:1162-1180):
   adds |d_a|^2 / 2 to *f and, if dchi is given, writes the upstream gradient
   conj(d_a) at c(a) of dchi, zero elsewhere */
static void blockenc_residual(int L, const ocplx* wchi, const ocplx* hpsi, double* f, ocplx* dchi)
{
	const int half = L / 2;
	const long s = 1L << half, n = 1L << L;
	if (dchi != NULL) { memset(dchi, 0, n * sizeof(ocplx)); }
	for (long a = 0; a < s; a++) {
		const long c = blockenc_embed(a, half);
		const ocplx d = wchi[c] - hpsi[a];
		*f += 0.5 * (creal(d) * creal(d) + cimag(d) * cimag(d));
		if (dchi != NULL) { dchi[c] = conj(d); }
	}
}

/* reference src/target.c:1100-1190 */
int orc_brickwall_blockenc_target(orc_linear_func hfunc, void* hdata, const ocplx* vlist, int nlayers, int L,
	const int* perms, double* fval)
{
	if (L % 2 != 0) { return -3; }
	const long s = 1L << (L / 2), n = 1L << L;
	ocplx* psi = xalloc(s * sizeof(ocplx));
	ocplx* hpsi = xalloc(s * sizeof(ocplx));
	ocplx* chi = xalloc(n * sizeof(ocplx));
	ocplx* wchi = xalloc(n * sizeof(ocplx));
	double f = 0;
	int ret = 0;
	for (long b = 0; b < s; b++) {
		if (blockenc_column(hfunc, hdata, L, b, psi, hpsi, chi) < 0) { ret = -2; break; }
		orc_apply_brickwall_unitary(vlist, nlayers, perms, L, chi, wchi);
		blockenc_residual(L, wchi, hpsi, &f, NULL);
	}
	*fval = f;
	free(wchi); free(chi); free(hpsi); free(psi);
	return ret;
}

/* reference src/target.c:1202-1330: per unit vector, forward with cache, the
   residual as upstream gradient, backward; gate gradients summed over e_b */
int orc_brickwall_blockenc_target_and_gradient(orc_linear_func hfunc, void* hdata, const ocplx* vlist, int nlayers,
	int L, const int* perms, double* fval, ocplx* dvlist)
{
	if (L % 2 != 0) { return -3; }
	const long s = 1L << (L / 2), n = 1L << L;
	const int ng = nlayers * (L / 2);
	ocplx* psi = xalloc(s * sizeof(ocplx));
	ocplx* hpsi = xalloc(s * sizeof(ocplx));
	ocplx* chi = xalloc(n * sizeof(ocplx));
	ocplx* wchi = xalloc(n * sizeof(ocplx));
	ocplx* cache = xalloc((size_t)ng * n * sizeof(ocplx));
	ocplx* dvu = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] = 0; }
	double f = 0;
	int ret = 0;
	for (long b = 0; b < s; b++) {
		if (blockenc_column(hfunc, hdata, L, b, psi, hpsi, chi) < 0) { ret = -2; break; }
		orc_brickwall_unitary_forward(vlist, nlayers, perms, L, chi, cache, wchi);
		blockenc_residual(L, wchi, hpsi, &f, chi);
		orc_brickwall_unitary_backward(vlist, nlayers, perms, L, cache, chi, wchi, dvu);
		for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] += dvu[e]; }
	}
	*fval = f;
	free(dvu); free(cache); free(wchi); free(chi); free(hpsi); free(psi);
	return ret;
}

/* reference src/target.c:1341-1365: conjugated gradient in tangent coordinates */
int orc_brickwall_blockenc_target_and_gradient_vector(orc_linear_func hfunc, void* hdata, const ocplx* vlist,
	int nlayers, int L, const int* perms, double* fval, double* grad_vec)
{
	ocplx* dv = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	const int ret = orc_brickwall_blockenc_target_and_gradient(hfunc, hdata, vlist, nlayers, L, perms, fval, dv);
	if (ret == 0) {
		for (int i = 0; i < nlayers; i++) {
			for (int e = 0; e < 16; e++) { dv[16*i + e] = conj(dv[16*i + e]); }
			orc_tangent_to_real(&vlist[16*i], &dv[16*i], &grad_vec[16*i]);
		}
	}
	free(dv);
	return ret;
}

/* reference src/brickwall_circuit.c:264-310: the placeholder states of every
   gate of layer l (from the cached state before it), each carried through all
   later gates, summed; psi_out is n x 16 interleaved */
int orc_apply_brickwall_unitary_gate_placeholder(const ocplx* vlist, int nlayers, const int* perms, int L, int l,
	const ocplx* cache, ocplx* psi_out)
{
	const long n = 1L << L;
	const int per = L / 2;
	ocplx* t[2] = { xalloc((size_t)n * 16 * sizeof(ocplx)), xalloc((size_t)n * 16 * sizeof(ocplx)) };
	memset(psi_out, 0, (size_t)n * 16 * sizeof(ocplx));
	for (int s2 = 0; s2 < L; s2 += 2) {
		const int g = l * per + s2 / 2;
		orc_apply_gate_placeholder(perms[l*L + s2], perms[l*L + s2 + 1], L, &cache[(size_t)g * n], t[0]);
		int cur = 0;
		for (int i = l; i < nlayers; i++) {
			for (int j = 0; j < L; j += 2) {
				if (i == l && j <= s2) { continue; }
				orc_apply_gate_to_array(&vlist[16*i], perms[i*L + j], perms[i*L + j + 1], L, 16, t[cur], t[1 - cur]);
				cur = 1 - cur;
			}
		}
		for (size_t e = 0; e < (size_t)n * 16; e++) { psi_out[e] += t[cur][e]; }
	}
	free(t[1]); free(t[0]);
	return 0;
}

/* reference src/target.c:1380-1620: hess1 from the backward Hessian pass with the
   residual as upstream gradient; hess2 = sum_a conj(P^dagger J_l) (P^dagger J_k)
   of the per-layer placeholder states (lower blocks k <= l computed, the upper
   ones their conjugate transposes); both summed over the unit vectors e_b */
int orc_brickwall_blockenc_target_gradient_hessian(orc_linear_func hfunc, void* hdata, const ocplx* vlist, int nlayers,
	int L, const int* perms, double* fval, ocplx* dvlist, ocplx* hess1, ocplx* hess2)
{
	if (L % 2 != 0) { return -3; }
	const int half = L / 2;
	const long s = 1L << half, n = 1L << L;
	const int ng = nlayers * half, m = 16 * nlayers;
	ocplx* psi = xalloc(s * sizeof(ocplx));
	ocplx* hpsi = xalloc(s * sizeof(ocplx));
	ocplx* chi = xalloc(n * sizeof(ocplx));
	ocplx* wchi = xalloc(n * sizeof(ocplx));
	ocplx* cache = xalloc((size_t)ng * n * sizeof(ocplx));
	ocplx* tmp = xalloc((size_t)n * 16 * sizeof(ocplx));
	ocplx* phi = xalloc((size_t)nlayers * s * 16 * sizeof(ocplx));
	ocplx* dvu = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	ocplx* hu1 = xalloc((size_t)m * m * sizeof(ocplx));
	ocplx* hu2 = xalloc((size_t)m * m * sizeof(ocplx));
	for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] = 0; }
	memset(hess1, 0, (size_t)m * m * sizeof(ocplx));
	memset(hess2, 0, (size_t)m * m * sizeof(ocplx));
	double f = 0;
	int ret = 0;
	for (long b = 0; b < s; b++) {
		if (blockenc_column(hfunc, hdata, L, b, psi, hpsi, chi) < 0) { ret = -2; break; }
		orc_brickwall_unitary_forward(vlist, nlayers, perms, L, chi, cache, wchi);
		blockenc_residual(L, wchi, hpsi, &f, chi);
		orc_brickwall_unitary_backward_hessian(vlist, nlayers, perms, L, cache, chi, wchi, dvu, hu1);
		for (int l = 0; l < nlayers; l++) {
			orc_apply_brickwall_unitary_gate_placeholder(vlist, nlayers, perms, L, l, cache, tmp);
			for (long a = 0; a < s; a++) {
				const long c = blockenc_embed(a, half);
				for (int j = 0; j < 16; j++) { phi[((size_t)l * s + a) * 16 + j] = tmp[c * 16 + j]; }
			}
		}
		memset(hu2, 0, (size_t)m * m * sizeof(ocplx));
		for (int l = 0; l < nlayers; l++) {
			for (int k = 0; k <= l; k++) {
				for (long a = 0; a < s; a++) {
					const ocplx* pl = &phi[((size_t)l * s + a) * 16];
					const ocplx* pk = &phi[((size_t)k * s + a) * 16];
					for (int i = 0; i < 16; i++) {
						for (int j = 0; j < 16; j++) { hu2[(size_t)(l*16 + i) * m + k*16 + j] += conj(pl[i]) * pk[j]; }
					}
				}
			}
		}
		for (int l = 0; l < nlayers; l++) {
			for (int k = l + 1; k < nlayers; k++) {
				for (int i = 0; i < 16; i++) {
					for (int j = 0; j < 16; j++) {
						hu2[(size_t)(l*16 + i) * m + k*16 + j] = conj(hu2[(size_t)(k*16 + j) * m + l*16 + i]);
					}
				}
			}
		}
		for (int e = 0; e < 16 * nlayers; e++) { dvlist[e] += dvu[e]; }
		for (size_t e = 0; e < (size_t)m * m; e++) { hess1[e] += hu1[e]; hess2[e] += hu2[e]; }
	}
	*fval = f;
	free(hu2); free(hu1); free(dvu); free(phi); free(tmp); free(cache); free(wchi); free(chi); free(hpsi); free(psi);
	return ret;
}

/* reference src/target.c:1631-1720 */
int orc_brickwall_blockenc_target_gradient_vector_hessian_matrix(orc_linear_func hfunc, void* hdata, const ocplx* vlist,
	int nlayers, int L, const int* perms, double* fval, double* grad_vec, double* H)
{
	const int m = 16 * nlayers;
	ocplx* dv = xalloc((size_t)nlayers * 16 * sizeof(ocplx));
	ocplx* hess1 = xalloc((size_t)m * m * sizeof(ocplx));
	ocplx* hess2 = xalloc((size_t)m * m * sizeof(ocplx));
	const int ret = orc_brickwall_blockenc_target_gradient_hessian(hfunc, hdata, vlist, nlayers, L, perms, fval, dv, hess1,
		hess2);
	if (ret == 0) {
		for (int i = 0; i < nlayers; i++) {
			for (int e = 0; e < 16; e++) { dv[16*i + e] = conj(dv[16*i + e]); }
			orc_tangent_to_real(&vlist[16*i], &dv[16*i], &grad_vec[16*i]);
		}
		project_hessian_matrix(nlayers, vlist, dv, hess1, hess2, H);
	}
	free(hess2); free(hess1); free(dv);
	return ret;
}
