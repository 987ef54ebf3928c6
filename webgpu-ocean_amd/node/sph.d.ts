// Type declarations of sph.js, matching sph/sph.ts:13-371.
export declare const sphParticleStructSize: 64;
export declare const posvelStructSize: 32;

export interface SPHParams {
    kernelRadius?: number;
    mass?: number;
    dt?: number;
    stiffness?: number;
    nearStiffness?: number;
    restDensity?: number;
    viscosity?: number;
}

export declare class SPHSimulator {
    numParticles: number;
    gridCount: number;
    kernelRadius: number;
    renderDiameter: number;
    particleBuffer: ArrayBuffer | null;
    posvelBuffer: ArrayBuffer | ArrayBufferView | null;
    constructor(particleBuffer: ArrayBuffer | null, posvelBuffer: ArrayBuffer | ArrayBufferView | null,
                renderDiameter: number, device?: { params?: SPHParams; random?: () => number });
    initDambreak(initHalfBoxSize: number[], numParticles: number): ArrayBuffer;
    reset(numParticles: number, initHalfBoxSize: number[], particles?: ArrayBuffer | ArrayBufferView): void;
    execute(commandEncoder?: unknown): void;
    changeBoxSize(realBoxSize: number[]): void;
    readPositions(dst?: ArrayBuffer | ArrayBufferView): ArrayBuffer | ArrayBufferView;
    posvelDevicePtr(): bigint;
    readParticles(dst?: ArrayBuffer | ArrayBufferView): ArrayBuffer | ArrayBufferView;
    writeParticles(src: ArrayBuffer | ArrayBufferView): void;
    synchronize(): void;
    destroy(): void;
}
