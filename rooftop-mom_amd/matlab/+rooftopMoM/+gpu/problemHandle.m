function handle = problemHandle(problem)
% problemHandle: Returns a new GPU problem (librooftop rt_problem_create) of a rooftopMoM.Problem: the same grid, face
%   sizes, wavenumber and quadrature orders (Problem.m:25-92, AssemblerContext.m:34). The generators are assembled on
%   the GPU when a product first needs them. The caller frees the handle with rooftop_mex('free', 'problem', handle);
%   a product made from it keeps the problem alive on its own.

    arguments
        problem (1, 1) rooftopMoM.Problem
    end

    [countX, countY] = problem.mesh().count();
    [sizeX, sizeY] = problem.mesh().faceSize();
    orders = problem.context().quadratureOrders();
    handle = rooftop_mex('problem', rooftopMoM.gpu.context(), countX, countY, sizeX, sizeY, problem.wavenumber(), ...
        double(orders), []);
end
