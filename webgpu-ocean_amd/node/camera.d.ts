// Type declarations of camera.js, matching camera.ts:4-101 and common.ts:1-9.
export declare const renderUniformsValues: ArrayBuffer;
export declare const renderUniformsViews: {
    texel_size: Float32Array;
    sphere_size: Float32Array;
    inv_projection_matrix: Float32Array;
    projection_matrix: Float32Array;
    view_matrix: Float32Array;
    inv_view_matrix: Float32Array;
};

type CanvasSize = { clientWidth: number; clientHeight: number } | { width: number; height: number };

export declare class Camera {
    currentXtheta: number;
    currentYtheta: number;
    currentDistance: number;
    target: number[];
    fov: number;
    zoomRate: number;
    constructor(canvasElement?: CanvasSize | null);
    reset(canvasElement: CanvasSize, initDistance: number, target: number[], fov: number,
          zoomRate: number): void;
    drag(deltaX: number, deltaY: number): void;
    wheel(deltaY: number): void;
    recalculateView(): void;
}

export declare function perspective(fov: number, aspect: number, near: number, far: number): number[];
export declare function lookAt(eye: number[], target: number[], up: number[]): number[];
export declare function inverse(m: number[]): number[];
