'use strict';
// Headless JavaScript host of the molecule.js compute-pass contract on the
// MI355X solver (librq.so through rq_addon.node).
//
// Configuration keys are the molecule.js page globals (molecule.js:4-152):
//   USER_NN, USER_SCREEN, USER_ATOMS [{i, j, k?, Z, el}], USER_Z, USER_RC,
//   USER_DV, USER_STEPS_PER_FRAME, USER_POISSON_INTERVAL, USER_JACOBI_PER_STEP,
//   USER_CURV_REG, USER_N_MOVE, USER_DT_NUC, USER_DAMPING, USER_FORCE_SCALE,
//   USER_DYNAMICS, FREEZE_BOUNDARY, NO_VEE, NO_NUC_REPULSION,
//   CONVERGENCE_THRESHOLD, USER_POTHER_FRAME, USER_SIC_INTERVAL,
//   USER_DIRECT_POTHER, USER_Z_NUC, USER_INIT_ZEFF / INIT_ZEFF,
//   USER_INIT_RCUT / INIT_RCUT, USER_SPLIT_TYPE / _IDX / _AXIS / _ROT.
// Host API mirrored: initGPU, doSteps(n) (molecule.js:2822-3261),
// restartWithPositions (:2127), activateAtoms (:2151), isGpuReady (:2196),
// onSweepDone (:4653).
//
// Not reproduced (outside the grid-solver path): USER_SPLIT_GROUPS frame
// tracking (molecule.js:1840-1875; the split axes stay as given), the
// NN-keyed hydrogen energy correction, thermostats, shear and restraints.

const path = require('path');
const addon = require(path.join(__dirname, 'rq_addon.node'));

function nucMass(z) { return ({1: 1, 2: 16, 3: 14, 4: 12}[z] || 1) * 1836; }

class MoleculeHost {
  constructor(cfg) {
    this.cfg = cfg;
    this.NN = cfg.USER_NN || 200;
    this.S = this.NN + 1;
    this.N2 = Math.round(this.NN / 2);
    this.screen = cfg.USER_SCREEN || 10;
    this.h = this.screen / this.NN;
    const atoms = cfg.USER_ATOMS;
    const Zs = cfg.USER_Z || atoms.map(a => a.Z);
    const rcs = cfg.USER_RC || atoms.map(a => a.rc || 0);
    this.nAtoms = atoms.length;
    this.Z = atoms.map((a, n) => Zs[n] || 0);
    this.Zorig = [...this.Z];
    this.activeCount = this.nAtoms;
    this.rc = atoms.map((a, n) => rcs[n] || 0);
    this.nucPos = atoms.map(a => [a.i, a.j, a.k !== undefined ? a.k : this.N2]);
    this.molPos = this.nucPos.map(p => [...p]);  // the molecule's positions (restart / activateAtoms)
    // nuclear charges (molecule.js:22-25): Z_nuc > 0 with Z = 0 is a bare nucleus
    this.Znuc = atoms.map((a, n) => (cfg.USER_Z_NUC ? cfg.USER_Z_NUC[n] || 0 : this.Z[n]));
    this.NELEC = this.Z.filter((z, n) => z > 0 || this.Znuc[n] > 0).length || 3;   // molecule.js:12-17
    this.N_ELECTRONS = this.Z.reduce((s, z) => s + z, 0);
    this.uniqueNuclei = this._uniqueNuclei(atoms);
    this.dv = cfg.USER_DV || (this.NELEC > 500 ? 0.01 : this.NELEC > 100 ? 0.03 : 0.12);
    this.STEPS_PER_FRAME = cfg.USER_STEPS_PER_FRAME ||
      (this.NELEC <= 30 ? 50 : this.NELEC <= 100 ? 5 : 2);
    this.POISSON_INTERVAL = cfg.USER_POISSON_INTERVAL || 50;
    this.JACOBI_PER_STEP = cfg.USER_JACOBI_PER_STEP || 2;
    this.NORM_INTERVAL = 20;
    this.N_MOVE = cfg.USER_N_MOVE || 10;
    this.DT_NUC = cfg.USER_DT_NUC || (this.NELEC <= 5 ? 2.0 : this.NELEC > 200 ? 0.2 : 0.8);
    this.DAMPING = cfg.USER_DAMPING || 0.98;
    this.MAX_VEL = this.NELEC <= 5 ? 0.3 : this.NELEC > 200 ? 0.03 : 0.1;
    this.forceScale = cfg.USER_FORCE_SCALE || 1.0;
    this.dynamicsEnabled = !!cfg.USER_DYNAMICS;
    this.curvReg = cfg.USER_CURV_REG !== undefined ? cfg.USER_CURV_REG : 0.15;
    this.threshold = cfg.CONVERGENCE_THRESHOLD || 0;
    // frame potential (PotherBuf) with self-interaction correction, molecule.js:112, 3009-3076
    this.POTHER_FRAME = cfg.USER_POTHER_FRAME !== undefined ? !!cfg.USER_POTHER_FRAME : true;
    this.SIC_INTERVAL = cfg.USER_SIC_INTERVAL !== undefined ? cfg.USER_SIC_INTERVAL :
      (this.NELEC <= 15 ? 1 : this.NELEC <= 30 ? 5 : 0);
    // per-electron direct Pother, the page's default up to 5 electrons (molecule.js:111)
    this.USE_DIRECT_POTHER = cfg.USER_DIRECT_POTHER !== undefined ? !!cfg.USER_DIRECT_POTHER : this.NELEC <= 5;
    this.addNucRepulsion = !cfg.NO_NUC_REPULSION;
    this.nucVel = this.nucPos.map(() => [0, 0, 0]);
    this.nucForce = this.nucPos.map(() => [0, 0, 0]);
    this.tStep = 0;
    this.frameCount = 0;
    this.E = 0; this.E_T = 0; this.E_eK = 0; this.E_ee = 0; this.E_KK = 0;
    this.dipole = [0, 0, 0];
    this.onSweepDone = null;
    this._stable = 0;
    this._lastE = Infinity;
    this.handle = null;
  }

  _atomRows() {
    return this.nucPos.map((p, n) => [p[0], p[1], p[2], this.Z[n], this.rc[n]]);
  }

  // Electrons whose nuclei share a start position share one nucleus of
  // charge sum(Z_nuc || Z) in E_KK and the nucleus-nucleus forces; the list
  // is fixed at start, like the page's (molecule.js:50-67).
  _uniqueNuclei(atoms) {
    const map = new Map();
    atoms.forEach((a, e) => {
      if (this.Z[e] === 0 && this.Znuc[e] === 0) return;
      const key = a.i + ',' + a.j + ',' + (a.k !== undefined ? a.k : this.N2);
      const q = this.Znuc[e] || this.Z[e];
      if (map.has(key)) { map.get(key).Zeff += q; map.get(key).elec.push(e); }
      else map.set(key, { Zeff: q, elec: [e] });
    });
    return [...map.values()];
  }

  // molecule.js per-atom fields (molecule.js:1875-1903) for the S2 init, K,
  // level set and forces; null when the page would use the defaults
  _atomsS2() {
    const c = this.cfg;
    const keys = ['USER_Z_NUC', 'USER_INIT_ZEFF', 'INIT_ZEFF', 'USER_INIT_RCUT', 'INIT_RCUT', 'USER_SPLIT_TYPE'];
    if (!keys.some(k => c[k] !== undefined)) return null;
    const split = c.USER_SPLIT_TYPE || [], idx = c.USER_SPLIT_IDX || [];
    const axis = c.USER_SPLIT_AXIS || [], rot = c.USER_SPLIT_ROT || [];
    return this.Z.map((z, n) => ({
      Z_nuc: this.Znuc[n],
      initZeff: c.USER_INIT_ZEFF ? c.USER_INIT_ZEFF[n] || z : (c.INIT_ZEFF || z),
      initRcut: c.USER_INIT_RCUT ? c.USER_INIT_RCUT[n] || 1e6 : (c.INIT_RCUT || 1e6),
      splitType: split[n] || 0, splitIdx: idx[n] || 0,
      splitAxis: axis[n] || [0, 0, 1], splitRot: rot[n] || 0,
    }));
  }

  _potherMode() {
    if (this._directPother()) return 'directSweeps';
    return this.POTHER_FRAME ? 'frame' : 'total';
  }

  initGPU() {
    this.handle = addon.create({
      NN: this.NN, screen: this.screen, dv: this.dv,
      poisson: 'twogrid', jacobiSweeps: this.JACOBI_PER_STEP, smoothA: 0.3333, smoothC: 9,
      twogridInterval: this.POISSON_INTERVAL, twogridCoarseSweeps: 10,
      boundary: this.cfg.FREEZE_BOUNDARY ? 'off' : 'levelset', curvReg: this.curvReg, dtW: 0.5,
      init: 's2', energy: 's2', rcDirichlet: 1,
      pother: this._potherMode(),
      sicInterval: this._potherMode() === 'frame' ? this.SIC_INTERVAL : 0,
    }, this._atomRows());
    const ext = this._atomsS2();
    if (ext) addon.setAtomsS2(this.handle, ext);
    addon.initTrial(this.handle);
    if (this._framePotential()) addon.potherFrame(this.handle, 0);  // PotherBuf = P (molecule.js:2091)
    return this.handle;
  }

  isGpuReady() { return this.handle !== null; }

  // the page's per-electron Poisson path: more than one electron, V_ee on (molecule.js:2833)
  _directPother() { return this.USE_DIRECT_POTHER && !this.cfg.NO_VEE && this.N_ELECTRONS > 1; }
  // PotherBuf stays zero with one electron or without V_ee (molecule.js:3010-3011)
  _framePotential() {
    return !this._directPother() && this.POTHER_FRAME && !this.cfg.NO_VEE && this.NELEC > 1;
  }

  // One frame of molecule.js doSteps (total-Poisson path).
  doSteps(n) {
    const h = this.handle;
    const t0 = Date.now();
    let forcesDue = false;
    for (let s = 0; s < n; s++) {
      if (this._directPother()) {
        addon.poisson(h, 1);                          // 10 damped sweeps of each P_m, Pother = P_m on m
      } else if (!this.cfg.NO_VEE) {
        addon.poisson(h, 1);                          // rho + JACOBI_PER_STEP damped sweeps
        if (s > 0 && s % this.POISSON_INTERVAL === 0) addon.twogridCorrect(h);
      }
      addon.update(h);
      if ((s + 1) % this.NORM_INTERVAL === 0 || s === n - 1) addon.normalize(h);
      if ((this.tStep + s + 1) % this.N_MOVE === 0) forcesDue = true;
    }
    // end-of-frame potential, counted from 1 like the page's draw loop
    if (this._framePotential()) addon.potherFrame(h, this.frameCount + 1);
    const e = addon.energy(h);
    if (!this.cfg.FREEZE_BOUNDARY) {
      for (let w = 0; w < Math.max(1, this.STEPS_PER_FRAME); w++) addon.boundary(h);
    }
    this.E_T = e.T; this.E_eK = e.VeK; this.E_ee = e.Vee;
    this.dipole = e.dipole;
    this.E_KK = this.nuclearEnergy();
    this.E = this.E_T + this.E_eK + this.E_ee + this.E_KK;
    if (!Number.isFinite(this.E)) throw new Error('Numerical instability at step ' + this.tStep);
    this.tStep += n;
    this.frameCount++;
    this.msPerStep = (Date.now() - t0) / n;
    if (forcesDue && this.dynamicsEnabled) this.moveNuclei(addon.forcesGradP(h, 0));
    this._checkConvergence();
    return this.E;
  }

  // E_KK over the unique nuclei, each at its first electron's position,
  // with the 0.04 h^2 softening of molecule.js:3166-3182.
  nuclearEnergy() {
    if (!this.addNucRepulsion) return 0;
    let e = 0;
    const h = this.h, U = this.uniqueNuclei;
    for (let a = 0; a < U.length; a++) {
      const pa = this.nucPos[U[a].elec[0]];
      for (let b = a + 1; b < U.length; b++) {
        const pb = this.nucPos[U[b].elec[0]];
        const dx = (pa[0] - pb[0]) * h, dy = (pa[1] - pb[1]) * h, dz = (pa[2] - pb[2]) * h;
        e += U[a].Zeff * U[b].Zeff / Math.sqrt(dx * dx + dy * dy + dz * dz + 0.04 * h * h);
      }
    }
    return e;
  }

  // Nucleus-nucleus Coulomb forces between unique nuclei (h^2 softening,
  // molecule.js:3226-3246), given to every electron of each nucleus.
  nuclearForces() {
    const f = this.nucPos.map(() => [0, 0, 0]);
    if (!this.addNucRepulsion) return f;
    const h = this.h, U = this.uniqueNuclei;
    for (let a = 0; a < U.length; a++) {
      const pa = this.nucPos[U[a].elec[0]];
      const fa = [0, 0, 0];
      for (let b = 0; b < U.length; b++) {
        if (b === a) continue;
        const pb = this.nucPos[U[b].elec[0]];
        const d = [0, 1, 2].map(c => (pa[c] - pb[c]) * h);
        const r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + h * h;
        const w = U[a].Zeff * U[b].Zeff / (Math.sqrt(r2) * r2);
        for (let c = 0; c < 3; c++) fa[c] += w * d[c];
      }
      for (const e of U[a].elec) f[e] = [...fa];
    }
    return f;
  }

  // Damped explicit nuclear step (molecule.js:4486-4492) on the electronic
  // force plus the unique-nucleus repulsion (molecule.js:3218-3254).
  moveNuclei(felec) {
    const h = this.h;
    const fnuc = this.nuclearForces();
    for (let a = 0; a < this.nAtoms; a++) {
      if (this.Z[a] === 0 && this.Znuc[a] === 0) continue;
      const f = [0, 1, 2].map(c => felec[3 * a + c] + fnuc[a][c]);
      this.nucForce[a] = f;
      const m = nucMass(this.Znuc[a] || this.Z[a]);
      for (let c = 0; c < 3; c++) {
        let v = this.nucVel[a][c] + f[c] / m * this.DT_NUC * this.forceScale;
        v *= this.DAMPING;
        v = Math.max(-this.MAX_VEL, Math.min(this.MAX_VEL, v));
        this.nucVel[a][c] = v;
        this.nucPos[a][c] = Math.max(5, Math.min(this.NN - 5, this.nucPos[a][c] + v * this.DT_NUC / h));
      }
    }
    addon.setAtoms(this.handle, this._atomRows());
    addon.recomputeK(this.handle);
  }

  restartWithPositions(pos) {
    for (let n = 0; n < pos.length && n < this.nAtoms; n++) {
      this.nucPos[n] = [pos[n][0], pos[n][1], pos[n][2]];
      this.molPos[n] = [...this.nucPos[n]];
    }
    this.nucVel = this.nucPos.map(() => [0, 0, 0]);
    this._restart();
  }

  // molecule.js:2150-2195: keep the first `count` nuclei charged (the rest
  // Z = 0), put the nuclei back at the molecule's positions and restart
  activateAtoms(count) {
    count = Math.max(1, Math.min(count, this.nAtoms));
    this.Z = this.Zorig.map((z, n) => (n < count ? z : 0));
    this.nucPos = this.molPos.map(p => [...p]);
    this.nucVel = this.nucPos.map(() => [0, 0, 0]);
    this.nucForce = this.nucPos.map(() => [0, 0, 0]);
    this.activeCount = count;
    this._restart();
  }

  _restart() {
    addon.setAtoms(this.handle, this._atomRows());
    addon.initTrial(this.handle);
    if (this._framePotential()) addon.potherFrame(this.handle, 0);
    this.tStep = 0;
    this.frameCount = 0;
    this._stable = 0;
    this._lastE = Infinity;
  }

  // Convergence rule of molecule.js:4644-4659: three consecutive frames with
  // |dE| below the threshold fire onSweepDone(E, steps, E_T, E_eK, E_ee, E_KK).
  _checkConvergence() {
    if (!this.threshold) return;
    if (Math.abs(this.E - this._lastE) < this.threshold) this._stable++;
    else this._stable = 0;
    this._lastE = this.E;
    if (this._stable >= 3 && this.onSweepDone) {
      const cb = this.onSweepDone;
      this.onSweepDone = null;
      cb(this.E, this.tStep, this.E_T, this.E_eK, this.E_ee, this.E_KK);
    }
  }

  readField(name) { return addon.readField(this.handle, name); }

  // molecule.js extract + slice readback (molecule.js:1108-1192, 1818): the
  // 3 S^2 + 2 S float buffer the page's renderer draws
  extractSlice(sliceK) {
    return addon.displaySlice(this.handle, sliceK === undefined ? this.N2 : sliceK);
  }

  destroy() {
    if (this.handle) addon.destroy(this.handle);
    this.handle = null;
  }
}

module.exports = { MoleculeHost, addon };
