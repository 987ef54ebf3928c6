// Writes sph_uniforms.json: the uniform values WebGPU-Ocean's SPHSimulator
// derives from its host constants (sph/sph.ts:41,58-77,140-174), evaluated
// with the JavaScript engine's own Math.pow / Math.ceil and rounded to f32 by
// Math.fround exactly where the reference stores them in Float32Array views.
// Run: node tests/golden/make_sph_uniforms.js > tests/golden/sph_uniforms.json
'use strict';
const f = Math.fround;
const kernelRadius = 0.07;
const cellSize = 1.0 * kernelRadius;
const halfMax = 2.0;
const len = 2.0 * halfMax;
const sentinel = 4 * cellSize;
const grids = Math.ceil((len + sentinel) / cellSize);
const out = {
  source: 'node ' + process.version,
  xGrids: grids,
  gridCount: grids * grids * grids,
  cellSize: f(cellSize),
  half: f(halfMax),
  offset: f(sentinel / 2),
  mass: f(1.0),
  kernelRadius: f(kernelRadius),
  kernelRadiusPow2: f(Math.pow(kernelRadius, 2)),
  kernelRadiusPow5: f(Math.pow(kernelRadius, 5)),
  kernelRadiusPow6: f(Math.pow(kernelRadius, 6)),
  kernelRadiusPow9: f(Math.pow(kernelRadius, 9)),
  dt: f(0.006),
  stiffness: f(20),
  nearStiffness: f(1.0),
  restDensity: f(15000),
  viscosity: f(100),
};
process.stdout.write(JSON.stringify(out, null, 1) + '\n');
