"""TEST INFRASTRUCTURE: checks that the oracle's pin tests can fail.

Each entry below plants one plausible mistake in a copy of oracle/pi_oracle.cpp
(a wrong index, sign, exponent, dropped term or halo plane, in QuotaMode,
CarbonMode, MacroMolecularMode, IronEnergyMode and the particle
interaction), builds the copy
into a temporary library, and runs the CPU pin tests against it
(PI_ORACLE_LIB).  Every mistake must make at least one pin fail.  The
committed source is never modified.

    python oracle/mutation_check.py            # all mistakes
    python oracle/mutation_check.py "z sign"   # one of them
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SRC = os.path.join(HERE, "pi_oracle.cpp")

BC_BOTTOM = """            if (bck[q][PI_BC_BOTTOM])
                for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
                    const int ii = i + Hx, jj = j + Hy, kk = Nz + Hz;
                    G[I(ii, jj, kk)] += getbc(q, PI_BC_BOTTOM, i, j, iter) * dt * Az(ii, jj) / vol(ii, jj, kk);
                }
"""
BC_TOP = """            if (bck[q][PI_BC_TOP])
                for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
                    const int ii = i + Hx, jj = j + Hy, kk = 1 + Hz;
                    G[I(ii, jj, kk)] -= getbc(q, PI_BC_TOP, i, j, iter) * dt * Az(ii, jj) / vol(ii, jj, kk);
                }
"""
BOTTOM_TOP, TOP_BOTTOM = BC_BOTTOM + BC_TOP, BC_TOP + BC_BOTTOM

INTERACT = """        for (const auto& pr : pint) particle_interaction(pr[0], pr[1]);   /* time_step.jl:120-128 */
        for (size_t q = 0; q < prel.size(); ++q) {                        /* :130-137 */
            pi_status st = particle_release(prel[q][0], prel[q][1], (int)q, dt);
            if (st != PI_OK) return st;
        }
        if (nsa > 0 && mode == PI_MODE_IRONENERGY)       /* diags_proc!(ptc, ptc, ac, ...) (:140-149) */
            for (int s = 0; s < nsp; ++s)
                for (int64_t i = 0; i < cap; ++i) diagSp[s][4][cell_of(sp[s], i)] += M::mask(sp[s].ac[i], sp[s].ptc[i]);
"""
NUM = """        for (int a = 0; a < nsa; ++a)                   /* diags_proc!(num, ac, ac, xi, yi, zi) (:152-159) */
            for (int64_t i = 0; i < cap; ++i) abNum[a][cell_of(ab[a], i)] += (FT)(ab[a].ac[i] ? 1 : 0);
"""
INTERACT_NUM, NUM_INTERACT = INTERACT + NUM, NUM + INTERACT

MISTAKES = [
    ("sinking dzC offset", "/ dzC_at(k) * dt;", "/ dzC_at(k + Hz) * dt;"),
    ("bottom flux_sink not zeroed at the last level", "planes_zero(c, 2, Nz + Hz, Hz + 1); }",
     "planes_zero(c, 2, Nz + Hz + 1, Hz); }"),
    ("z sign", "oz[i] = pz[i] - M::mask(ac[i], wt * w[i] * dt);", "oz[i] = pz[i] + M::mask(ac[i], wt * w[i] * dt);"),
    ("w face offset", "FT zs = comp == 2 ? z + L(1) : z + L(0.5f);", "FT zs = z + L(0.5f);"),
    ("PAR with dzC for dzF", "FT atten = (Chl[I(ii, jj, kk)] / vol(ii, jj, kk) * kc + kw) * dzF_at(kk);",
     "FT atten = (Chl[I(ii, jj, kk)] / vol(ii, jj, kk) * kc + kw) * dzC_at(kk);"),
    ("nitrification source", "*F[PI_NO3] = *F[PI_NO3] + T(PI_NH4, n) * Nit * dt;",
     "*F[PI_NO3] = *F[PI_NO3] + T(PI_NO3, n) * Nit * dt;"),
    ("dust scavenging dropped in PFe_inorg", "(lmin + ldust * T(PI_Dust, n)) * DFe * DFeFrac * dt +\n                               excess",
     "(lmin) * DFe * DFeFrac * dt +\n                               excess"),
    ("respiration sign", "p.Cq[i] = p.Cq[i] + (p.VDOC[i] - p.resp[i]) * dt;",
     "p.Cq[i] = p.Cq[i] + (p.VDOC[i] + p.resp[i]) * dt;"),
    ("DIC consumption sign", "plk[PI_DIC][c] += M::mask(a, (p.resp[i] - p.PS[i]) * dt);",
     "plk[PI_DIC][c] += M::mask(a, (p.PS[i] - p.resp[i]) * dt);"),
    ("exudation dropped", "plk[PI_DOC][c] += M::mask(a, (p.exu[i] - p.VDOC[i]) * dt);",
     "plk[PI_DOC][c] += M::mask(a, (- p.VDOC[i]) * dt);"),
    ("mortality with pow 2", "shape_dec(p.Sz[i], q[PI_P_mort_reg], L(1.0e-5f), 4)",
     "shape_dec(p.Sz[i], q[PI_P_mort_reg], L(1.0e-5f), 2)"),
    ("Sizer with pow 4", "shape_inc(Sz, q[PI_P_dvid_reg], L(2.0e-3f), 2)", "shape_inc(Sz, q[PI_P_dvid_reg], L(2.0e-3f), 4)"),
    ("Timer hour in minutes", "FT P_ = q[PI_P_dvid_P], Sz = p.Sz[i], v;\n            FT hour = M::rem(t, L(86400)) / L(3600);",
     "FT P_ = q[PI_P_dvid_P], Sz = p.Sz[i], v;\n            FT hour = M::rem(t, L(86400)) / L(60);"),
    ("all species' pop when shared_graz != 1",
     "                std::fill(pop.begin(), pop.end(), (FT)0);\n                acc_counts(sp[s]);",
     "                std::fill(pop.begin(), pop.end(), (FT)0);\n                for (int s2 = 0; s2 < nsp; ++s2) acc_counts(sp[s2]);"),
    ("bounded velocity halo like a tracer halo", "else planes_bcast(c, dim, N + H + 2, H - 1, N + H + 1);",
     "else planes_bcast(c, dim, N + H + 1, H, N + H);"),
    ("grazing loss without Cq",
     "plk[PI_POC][c] += M::mask(a, (p.Bm[i] + p.Cq[i]) * (L(1) - fC)) * loss[i];\n"
     "            plk[PI_PON][c] += M::mask(a, (p.Bm[i] * R_NC + p.Nq[i]) * (L(1) - fN)) * loss[i];",
     "plk[PI_POC][c] += M::mask(a, (p.Bm[i]) * (L(1) - fC)) * loss[i];\n"
     "            plk[PI_PON][c] += M::mask(a, (p.Bm[i] * R_NC + p.Nq[i]) * (L(1) - fN)) * loss[i];"),
    ("z diffusion sign", "FT ddz = M::mask(LM(i, j, k), M::mask(LM(i, j, k + 1), fz(i, j, k) - fz(i, j, k + 1)));",
     "FT ddz = M::mask(LM(i, j, k), M::mask(LM(i, j, k + 1), fz(i, j, k + 1) - fz(i, j, k)));"),
    ("z flux from the wrong side", "return L(0.5f) * (T + std::fabs(T)) * (c[I(i, j, k)] + pm * d) +",
     "return L(0.5f) * (T + std::fabs(T)) * (c[I(i, j, k - 1)] + pm * d) +"),
    ("x flux from the wrong side", "return L(0.5f) * (T + std::fabs(T)) * (c[I(i - 1, j, k)] + pp * d) +",
     "return L(0.5f) * (T + std::fabs(T)) * (c[I(i, j, k)] + pp * d) +"),
    ("division: daughter takes the first free slot",
     "/* get_tind_kernel! */\n            if (p.dvid[i] == L(1)) p.idx[i] = free_ind[(size_t)(idc_int[i] - 1)];",
     "/* get_tind_kernel! */\n            if (p.dvid[i] == L(1)) p.idx[i] = free_ind[0];"),
    ("PS without the Chl:C ratio", "FT PS = PCm * (one - M::exp(-aI / M::max(e30, PCm) * p.Chl[i] / M::max(e30, p.Bm[i])))",
     "FT PS = PCm * (one - M::exp(-aI / M::max(e30, PCm) / M::max(e30, p.Bm[i])))"),
    ("NO3 uptake with the NH4 half-saturation", "tNO3[i] / M::max(e30, tNO3[i] + q[PI_P_KsatNO3])",
     "tNO3[i] / M::max(e30, tNO3[i] + q[PI_P_KsatNH4])"),
    ("rhoChl light gate at 1", "M::max(e30, tpar[i] * q[PI_P_alpha] * q[PI_P_Phi] * p.Chl[i] / p.Bm[i]);\n"
     "            r = M::mask(M::isless(L(1.0e-1f), tpar[i]), r);",
     "M::max(e30, tpar[i] * q[PI_P_alpha] * q[PI_P_Phi] * p.Chl[i] / p.Bm[i]);\n"
     "            r = M::mask(M::isless(L(1.0f), tpar[i]), r);"),
    ("DOC uptake gate at 10 %", "M::isless(L(1.0e-2f), p.PS[i] / (p.VDOC[i] + p.PS[i]))",
     "M::isless(L(1.0e-1f), p.PS[i] / (p.VDOC[i] + p.PS[i]))"),
    ("tempFunc_PS shape constant", "FT x = T - q[PI_P_Topt], xmax = q[PI_P_Tmax] - q[PI_P_Topt];\n"
     "        FT regT = shape_dec(x, xmax, L(4.0e-2f), 4);",
     "FT x = T - q[PI_P_Topt], xmax = q[PI_P_Tmax] - q[PI_P_Topt];\n"
     "        FT regT = shape_dec(x, xmax, L(4.0e-3f), 4);"),
    ("gather NO3 from NH4", "tNO3[i] = M::mask(a, M::max(z0, tr[PI_NO3][c]));", "tNO3[i] = M::mask(a, M::max(z0, tr[PI_NH4][c]));"),
    ("uptake cap without pop", "FT pp = M::max(one, tpop[i]);", "FT pp = one;"),
    ("Chl synthesis without R_NC", "p.Chl[i] = p.Chl[i] + dt * p.BS[i] * q[PI_P_R_NC] * p.rho[i];",
     "p.Chl[i] = p.Chl[i] + dt * p.BS[i] * p.rho[i];"),
    ("find_inds without the +1", "p.xi[i] = (int64_t)M::mask(p.ac[i], p.x[i] + L(1)) + Hx;",
     "p.xi[i] = (int64_t)M::mask(p.ac[i], p.x[i]) + Hx;"),
    ("carbon: tempFunc_PS shape constant", "FT regT = shape_dec(x, xmax, L(1.0e-1f), 4);",
     "FT regT = shape_dec(x, xmax, L(4.0e-2f), 4);"),
    ("carbon: bacteria term dropped", "(light * (one - q[PI_C_is_bact]) + q[PI_C_is_bact])",
     "(light * (one - q[PI_C_is_bact]))"),
    ("carbon: thermal damage sign", "FT td = (T - q[PI_C_Topt]) * q[PI_C_f_T2B] * (Bm - Bd);",
     "FT td = (q[PI_C_Topt] - T) * q[PI_C_f_T2B] * (Bm - Bd);"),
    ("carbon: repair added to the damage", "p.Cq[i] = p.Cq[i] + (p.TD[i] - p.RP[i]) * dt;",
     "p.Cq[i] = p.Cq[i] + (p.TD[i] + p.RP[i]) * dt;"),
    ("carbon: daughter copies Bd", "p.Sz[d] = p.Sz[i]; p.Bm[d] = p.Bm[i]; p.Chl[d] = p.Chl[i];",
     "p.Sz[d] = p.Sz[i]; p.Bm[d] = p.Bm[i]; p.Chl[d] = p.Chl[i]; p.Cq[d] = p.Cq[i];"),
    ("carbon: DOC loss fraction swapped", "plk[PI_DOC][c] += M::mask(a, p.Bm[i] * fC) * loss[i];",
     "plk[PI_DOC][c] += M::mask(a, p.Bm[i] * (L(1) - fC)) * loss[i];"),
    ("carbon: thermal mortality threshold", "L(0.99f), L(1.0e-6f), 4)", "L(0.9f), L(1.0e-6f), 4)"),
    ("carbon: respiration with tempFunc_PS", "q[PI_C_respir] * p.Bm[i] * c_tempFunc(tT[i], q)",
     "q[PI_C_respir] * p.Bm[i] * c_tempFunc_PS(tT[i], q)"),
    ("carbon: Chl from Cquota", "p.Chl[i] = p.Bm[i] * q[PI_C_Chl2C];", "p.Chl[i] = p.Bm[i] * q[PI_C_Cquota];"),
    ("carbon: PS on the whole biomass", "M::max(z0, p.Bm[i] - p.Cq[i])", "M::max(z0, p.Bm[i])"),
    ("flux bc: east adds", "G[I(ii, jj, kk)] -= getbc(q, PI_BC_EAST, j, k, iter)",
     "G[I(ii, jj, kk)] += getbc(q, PI_BC_EAST, j, k, iter)"),
    ("flux bc: south with Ax", "getbc(q, PI_BC_SOUTH, i, k, iter) * dt * Ay(ii, jj, kk)",
     "getbc(q, PI_BC_SOUTH, i, k, iter) * dt * Ax(ii, jj, kk)"),
    ("flux bc: west on the halo plane", "const int ii = 1 + Hx, jj = j + Hy, kk = k + Hz;",
     "const int ii = Hx, jj = j + Hy, kk = k + Hz;"),
    ("flux bc: bottom on the top level", "const int ii = i + Hx, jj = j + Hy, kk = Nz + Hz;",
     "const int ii = i + Hx, jj = j + Hy, kk = 1 + Hz;"),
    ("flux bc: face array transposed", "return bca[q][f][(int64_t)(a - 1) + na * ((int64_t)(b - 1) + nb * tt)];",
     "return bca[q][f][(int64_t)(b - 1) + nb * ((int64_t)(a - 1) + na * tt)];"),
    ("flux bc: series read at the first iteration",
     "const int64_t tt = bck[q][f] == PI_BCK_SERIES ? (int64_t)(t - 1) : 0;", "const int64_t tt = 0;"),
    ("flux bc: top before bottom", BOTTOM_TOP, TOP_BOTTOM),
    ("latlon: corner divisor of the row below", "F(xi, yi + 1, zi) / D(xi, yi + 1, zi);",
     "F(xi, yi + 1, zi) / D(xi, yi, zi);"),
    ("latlon: walk without the second halo offset", "const int ci = clampi(p.xi[i] + Hx, Px), cj = clampi(p.yi[i] + Hy, Py);",
     "const int ci = clampi(p.xi[i], Px), cj = clampi(p.yi[i], Py);"),
    ("latlon: x transport with Ay", "FT T = Ax(i, j, k) * u[I(i, j, k)];", "FT T = Ay(i, j, k) * u[I(i, j, k)];"),
    ("latlon: y transport difference with one area",
     "Ay(i, j + 1, k) * w[I(i, j + 1, k)] - Ay(i, j, k) * w[I(i, j, k)]",
     "Ay(i, j, k) * w[I(i, j + 1, k)] - Ay(i, j, k) * w[I(i, j, k)]"),
    ("latlon: CFL with dxF", "FT C = std::fabs(u[I(i, j, k)] * dt / DxC(i, j));",
     "FT C = std::fabs(u[I(i, j, k)] * dt / DxF(i, j));"),
    ("latlon: meridional diffusion over dyC", "(dy0(c, i, j, k) / DyF(i, j))", "(dy0(c, i, j, k) / DyC(i, j))"),
    ("latlon: sweep over the volume of the cell north",
     "tt = tt - dt / vol(i, j, k) * (dG - c[I(i, j, k)] * dT);", "tt = tt - dt / vol(i, j + 1, k) * (dG - c[I(i, j, k)] * dT);"),
    ("latlon: PAR over the column area", "FT atten = (Chl[I(ii, jj, kk)] / vol(ii, jj, kk) * kc + kw) * dzF_at(kk);",
     "FT atten = (Chl[I(ii, jj, kk)] / (Az(ii, jj) * dzF_at(kk)) * kc + kw) * dzF_at(kk);"),
    ("latlon: exchange over the volume below", "tr[q][n] = tr[q][n] + (Gcs[q][n] + plk[q][n] / vol(i, j, k));",
     "tr[q][n] = tr[q][n] + (Gcs[q][n] + plk[q][n] / vol(i, j, k + 1));"),
    ("abiotic: created at the cell centre", "p.x[r] = (FT)(float)(i - Hx) - L(0.4f);",
     "p.x[r] = (FT)(float)(i - Hx) - L(0.5f);"),
    ("abiotic: draw of the next cell", "pi_uniform3(seed, iteration, PI_ABIO_SPECIES + a, PI_KIND_ABIO_BC, c, &u0, &u1, &u2);",
     "pi_uniform3(seed, iteration, PI_ABIO_SPECIES + a, PI_KIND_ABIO_BC, c + 1, &u0, &u1, &u2);"),
    ("abiotic: top flux through Ax", "put(i + Hx, j + Hy, 1 + Hz, bcv(PI_BC_TOP, i, j), Az(i + Hx, j + Hy));",
     "put(i + Hx, j + Hy, 1 + Hz, bcv(PI_BC_TOP, i, j), Ax(i + Hx, j + Hy, 1 + Hz));"),
    ("abiotic: free slots from the end", "for (int64_t i = 0; i < cap; ++i) if (!p.ac[i]) freei.push_back(i);",
     "for (int64_t i = cap - 1; i >= 0; --i) if (!p.ac[i]) freei.push_back(i);"),
    ("abiotic: walk on the phytoplankton sub-stream", "particle_diffusion(ab[a], PI_ABIO_SPECIES + a, dt);",
     "particle_diffusion(ab[a], a, dt);"),
    ("abiotic: num counts every slot", "abNum[a][cell_of(ab[a], i)] += (FT)(ab[a].ac[i] ? 1 : 0);",
     "abNum[a][cell_of(ab[a], i)] += (FT)1;"),
    ("abiotic: series one iteration late", "const int64_t tt = abck[a][f] == PI_BCK_SERIES ? iteration - 1 : 0;",
     "const int64_t tt = abck[a][f] == PI_BCK_SERIES ? iteration : 0;"),
    ("abiotic: the first face decides at an edge",
     "flag[c] = M::isless(rnd(c) * szN * Ns, b * dt * A) ? (FT)1 : (FT)0;",
     "if (flag[c] == (FT)0) flag[c] = M::isless(rnd(c) * szN * Ns, b * dt * A) ? (FT)1 : (FT)0;"),
    ("mm: protein synthesis from protein", "p.S_PRO[i] = q[PI_M_k_pro] * RNA[i] * limit_PRO * tf;",
     "p.S_PRO[i] = q[PI_M_k_pro] * PRO[i] * limit_PRO * tf;"),
    ("mm: synthesis masked by ac", "p.S_PRO[i] = q[PI_M_k_pro] * RNA[i] * limit_PRO * tf;",
     "p.S_PRO[i] = M::mask(p.ac[i], q[PI_M_k_pro] * RNA[i] * limit_PRO * tf);"),
    ("mm: protein respired with the DNA N:C", "NST[i] = NST[i] + M::max(z0, (z0 - CH[i])) * q[PI_M_R_NC_PRO];",
     "NST[i] = NST[i] + M::max(z0, (z0 - CH[i])) * q[PI_M_R_NC_DNA];"),
    ("mm: exudation taken times dt", "CH[i] = CH[i] - p.exu[i];", "CH[i] = CH[i] - p.exu[i] * dt;"),
    ("mm: DNA synthesis not gated",
     "p.S_DNA[i] = M::mask(M::isless(DNA[i] / (q[PI_M_C_DNA] * Ns), L(2)), q[PI_M_k_dna] * PRO[i] * limit_DNA * tf);",
     "p.S_DNA[i] = q[PI_M_k_dna] * PRO[i] * limit_DNA * tf;"),
    ("mm: mortality falling with age", "shape_inc(p.age[i], q[PI_M_mort_reg], L(1.0e-4f), 4)",
     "shape_dec(p.age[i], q[PI_M_mort_reg], L(1.0e-4f), 4)"),
    ("mm: division below two DNA units", "M::isless(L(2), DNA[i] / (q[PI_M_C_DNA] * q[PI_M_Nsuper]))",
     "M::isless(DNA[i] / (q[PI_M_C_DNA] * q[PI_M_Nsuper]), L(2))"),
    ("mm: total N without chlorophyll", "RNA * q[PI_M_R_NC_RNA] + NST + Chl / L(893.49f) * L(4.0f);",
     "RNA * q[PI_M_R_NC_RNA] + NST;"),
    ("mm: daughter without RNA",
     "PRO[d] = PRO[i]; DNA[d] = DNA[i]; RNA[d] = RNA[i]; CH[d] = CH[i];", "PRO[d] = PRO[i]; DNA[d] = DNA[i]; CH[d] = CH[i];"),
    ("iron: a single return at sunset", "FT f_PS2ST = q[PI_I_k_Fe_PS2ST] * qFePS * (FT)(falling + dark);",
     "FT f_PS2ST = q[PI_I_k_Fe_PS2ST] * qFePS * (FT)(falling || dark);"),
    ("iron: DIC charged with the carbon fixed", "plk[PI_DIC][c] += M::mask(a, (p.RS[i] - p.PS[i]) * dt);",
     "plk[PI_DIC][c] += M::mask(a, (p.RS[i] - p.CF[i]) * dt);"),
    ("iron: surface area as Sz^(1/3)", "M::pow(p.Sz[i], L(2.0f) / L(3.0f))", "M::pow(p.Sz[i], L(1.0f) / L(3.0f))"),
    ("iron: the NH4 reserve not returned", "plk[PI_NH4][c] += M::mask(a, p.Nq[i]) * loss[i];", "(void)0;"),
    ("iron: lost iron without the NR pool", "FT Fe = p.qFe[i] + p.qFePS[i] + p.qFeNF[i] + p.qFeNR[i];",
     "FT Fe = p.qFe[i] + p.qFePS[i] + p.qFeNF[i];"),
    ("iron: nitrate reduction adds to qNO3", "p.qNO3[i] = p.qNO3[i] + -p.NR[i] * dt;",
     "p.qNO3[i] = p.qNO3[i] + p.NR[i] * dt;"),
    ("iron: light does not reset tdark",
     "p.tdark[i] = p.tdark[i] * (one - (FT)M::isless(L(0), tpar[i])) + M::mask(M::isequal(L(0), tpar[i]), dt);",
     "p.tdark[i] = p.tdark[i] + M::mask(M::isequal(L(0), tpar[i]), dt);"),
    ("iron: dead slots keep their iron", "FT f_PS2ST = q[PI_I_k_Fe_PS2ST] * qFePS * (FT)(falling + dark);",
     "FT f_PS2ST = M::mask(p.ac[i], q[PI_I_k_Fe_PS2ST] * qFePS * (FT)(falling + dark));"),
    ("iron: storage to photosynthesis on falling light", "FT f_ST2PS = M::mask(rising, q[PI_I_k_Fe_ST2PS] * qFe);",
     "FT f_ST2PS = M::mask(falling, q[PI_I_k_Fe_ST2PS] * qFe);"),
    ("interaction: residue of the 0-based index", "const int64_t slot = ((i + 1) % mc) + 1;",
     "const int64_t slot = (i % mc) + 1;"),
    ("interaction: the first individual of a residue kept", "intac[(size_t)(slot - 1) + (size_t)mc * k] = i + 1;",
     "if (!intac[(size_t)(slot - 1) + (size_t)mc * k]) intac[(size_t)(slot - 1) + (size_t)mc * k] = i + 1;"),
    ("interaction: across cells", "if (xi1 == xi2 && yi1 == yi2 && zi1 == zi2) {", "if (true) {"),
    ("interaction: distance in grid units", "FT dx = (x1 - x2) * DxF((int)xi1, (int)yi1);", "FT dx = (x1 - x2);"),
    ("interaction: max_ptc inclusive", "if (id > 0 && pl.ptc[id - 1] < maxptc) {", "if (id > 0 && pl.ptc[id - 1] <= maxptc) {"),
    ("interaction: particles in reverse slot order",
     "for (int64_t k = 0; k < cap; ++k) {\n            if (!pa.ac[k]) continue;\n            for (int j = 1;",
     "for (int64_t k = cap - 1; k >= 0; --k) {\n            if (!pa.ac[k]) continue;\n            for (int j = 1;"),
    ("release: x from a fresh draw", "pa.x[k] = pl.x[i] + rx[i];", "pa.x[k] = pl.x[i] + (FT)b0;"),
    ("release: y from the first species' draws", "pi_uniform3(seed, iteration, nsp - 1, PI_KIND_DECISION, i, &b0, &b1, &b2);",
     "pi_uniform3(seed, iteration, 0, PI_KIND_DECISION, i, &b0, &b1, &b2);"),
    ("release: z with the mortality draw", "pa.z[k] = pl.z[i] + (FT)b2;", "pa.z[k] = pl.z[i] + (FT)b1;"),
    ("release: free slots from the end", "for (int64_t k = 0; k < cap; ++k) if (!pa.ac[k]) free_ind.push_back(k + 1);",
     "for (int64_t k = cap - 1; k >= 0; --k) if (!pa.ac[k]) free_ind.push_back(k + 1);"),
    ("release: ptc kept", "                pl.ptc[i] -= L(1);\n", "\n"),
    ("release: probability without ptc", "R[i] = (M::isless((FT)u0, rP * dt * pl.ptc[i]) && pl.ac[i]) ? L(1) : L(0);",
     "R[i] = (M::isless((FT)u0, rP * dt) && pl.ac[i]) ? L(1) : L(0);"),
    ("ptc diagnostic of every slot", "diagSp[s][4][cell_of(sp[s], i)] += M::mask(sp[s].ac[i], sp[s].ptc[i]);",
     "diagSp[s][4][cell_of(sp[s], i)] += sp[s].ptc[i];"),
    ("abiotic num before the interaction", INTERACT_NUM, NUM_INTERACT),
]


def main(only):
    orig = open(SRC).read()
    caught, missed = [], []
    with tempfile.TemporaryDirectory() as tmp:
        for name, a, b in MISTAKES:
            if only and name not in only:
                continue
            if orig.count(a) != 1:
                raise SystemExit(f"{name}: the text to change is not unique in pi_oracle.cpp")
            d = os.path.join(tmp, "oracle")
            os.makedirs(d, exist_ok=True)
            src = os.path.join(d, "pi_oracle.cpp")
            open(src, "w").write(orig.replace(a, b))
            lib = os.path.join(tmp, "liboracle_mut.so")
            subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                            "-fno-finite-math-only", "-fopenmp", "-fPIC", "-shared", "-iquote", HERE, "-o", lib, src], check=True)
            env = dict(os.environ, PI_ORACLE_LIB=lib)
            r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                                "tests/test_oracle_pins.py", "tests/test_oracle_pins_physics.py",
                                "tests/test_oracle_pins_carbon.py", "tests/test_oracle_pins_bcs.py",
                                "tests/test_oracle_pins_latlon.py", "tests/test_oracle_pins_abiotic.py",
                                "tests/test_oracle_pins_mm.py", "tests/test_oracle_pins_iron.py",
                                "tests/test_oracle_pins_interact.py"],
                               cwd=REPO, env=env, capture_output=True, text=True)
            first = next((l for l in r.stdout.splitlines() if l.startswith("FAILED")), "")
            (caught if r.returncode != 0 else missed).append(name)
            print(f"{'caught' if r.returncode else 'MISSED'}: {name}  {first}", flush=True)
    print(f"{len(caught)} caught, {len(missed)} missed", flush=True)
    return 1 if missed else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
