// Type declarations of mls-mpm.js, matching mls-mpm/mls-mpm.ts:10-276.
export declare const mlsmpmParticleStructSize: 80;
export declare const posvelStructSize: 32;

export interface MPMParams {
    stiffness?: number;
    restDensity?: number;
    dynamicViscosity?: number;
    dt?: number;
    fixedPointMultiplier?: number;
}

export declare class MLSMPMSimulator {
    numParticles: number;
    gridCount: number;
    renderDiameter: number;
    particleBuffer: ArrayBuffer | null;
    posvelBuffer: ArrayBuffer | ArrayBufferView | null;
    constructor(particleBuffer: ArrayBuffer | null, posvelBuffer: ArrayBuffer | ArrayBufferView | null,
                renderDiameter: number, device?: { params?: MPMParams; random?: () => number });
    initDambreak(initBoxSize: number[], numParticles: number): ArrayBuffer;
    reset(numParticles: number, initBoxSize: number[], particles?: ArrayBuffer | ArrayBufferView): void;
    execute(commandEncoder?: unknown): void;
    changeBoxSize(realBoxSize: number[]): void;
    readPositions(dst?: ArrayBuffer | ArrayBufferView): ArrayBuffer | ArrayBufferView;
    posvelDevicePtr(): bigint;
    readParticles(dst?: ArrayBuffer | ArrayBufferView): ArrayBuffer | ArrayBufferView;
    writeParticles(src: ArrayBuffer | ArrayBufferView): void;
    synchronize(): void;
    destroy(): void;
}
